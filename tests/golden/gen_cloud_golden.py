"""Write tests/golden/cloud_knn.npz from the compiled reference's PointCloudBVH<double>
(oracle/_ref, ebgref_cloud_knn): closestPoints of tests/cloud_cases.queries() for k = 1, 4, 16 and
allNearestNeighbors for k = 1, 8, on the cloud of tests/cloud_cases.cloud() (float32 points
promoted to double). Run where the reference sources are available:
``make -C oracle ref && python tests/golden/gen_cloud_golden.py``.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import cloud_cases as cc  # noqa: E402
from oracle import ref  # noqa: E402


def main():
    P, Q = cc.cloud(), cc.queries()
    out = {"cloud_sha256": np.frombuffer(hashlib.sha256(P.tobytes()).digest(), np.uint8),
           "queries_sha256": np.frombuffer(hashlib.sha256(Q.tobytes()).digest(), np.uint8)}
    for k in (1, 4, 16):
        i, d = ref.cloud_knn(P, Q, k)
        out[f"knn{k}.index"], out[f"knn{k}.d2"] = i.astype(np.int32), d
    for k in (1, 8):
        i, d = ref.cloud_knn(P, None, k, self_query=True)
        out[f"self{k}.index"], out[f"self{k}.d2"] = i.astype(np.int32), d
    np.savez_compressed(os.path.join(HERE, "cloud_knn.npz"), **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
