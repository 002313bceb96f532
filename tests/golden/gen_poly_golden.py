"""Write tests/golden/polymesh.npz from the compiled reference's MeshSDF<double, uint32_t, K> over
polygon faces (oracle/_ref, ebgref_polymesh_query): for every mesh of tests/poly_cases.py (float32
vertices promoted to double), the mesh itself and, at 3,000 query points (poly_cases.queries, seed
0), the signed distance and the first face of getClosestFaces(p, true) with its unsigned
distance. Run where the reference sources are available:
``make -C oracle ref && python tests/golden/gen_poly_golden.py``.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import poly_cases as pc  # noqa: E402
from oracle import ref  # noqa: E402


def main():
    out = {}
    for name in pc.NAMES:
        V, F = pc.mesh(name)
        off, idx = ref.flatten_faces(F)
        Q = pc.queries(V, 3000, seed=0)
        d, face, dist = ref.PolyMesh(V.astype(np.float64), F).query(Q.astype(np.float64))
        out.update({f"{name}.verts": V, f"{name}.offsets": off, f"{name}.indices": idx, f"{name}.queries": Q,
                    f"{name}.d": d, f"{name}.face": face, f"{name}.dist": dist})
    np.savez_compressed(os.path.join(HERE, "polymesh.npz"), **out)
    print({k: v.shape for k, v in out.items() if k.endswith(".d")})


if __name__ == "__main__":
    main()
