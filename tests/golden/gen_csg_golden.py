"""Write tests/golden/csg_algebra.npz from the compiled reference: for every expression of
tests/csg_cases.py, the reference's value() (built with its own constructors and factories through
oracle/_ref/libebgref.so, ebgref_expr_value) at 2,000 seeded query points.

Run where the reference sources are available (EBGEOMETRY_HOME, default /root/reference):
``make -C oracle ref && python tests/golden/gen_csg_golden.py``. The fixture holds data only: per
case the reference values and a checksum of the regenerated queries (csg_cases.queries).
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

import csg_cases as cc  # noqa: E402
from oracle import ref  # noqa: E402

N = 2000


def main():
    out = {}
    for name in cc.NAMES:
        f, meshes, lo, hi = cc.case(name)
        q = cc.queries(name, N)
        rm, idx = [], {}
        for k, (v, fa) in meshes.items():
            idx[k] = len(rm)
            rm.append(ref.Mesh(np.asarray(v, np.float64), fa))
        out[f"{name}.value"] = ref.expr_value(f, q.astype(np.float64), rm, idx)
        out[f"{name}.queries_sha256"] = np.frombuffer(hashlib.sha256(q.tobytes()).digest(), np.uint8)
        print(name, out[f"{name}.value"][:3])
    np.savez_compressed(os.path.join(HERE, "csg_algebra.npz"), **out)


if __name__ == "__main__":
    main()
