"""CPU-side checks of the C ABI: the library loads, exports every entry point of
include/ebgpu.h, and its host-only pseudo-normals equal the reference's."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ebgpu.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ebg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_function():
    from ebgeometry_amd import _lib
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers them
    covered = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) <= covered, set(names) - covered


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of the option structs have the C compiler's size and field offsets."""
    import subprocess
    from ebgeometry_amd import _lib
    structs = {"ebg_build_opts": _lib.BuildOpts, "ebg_query_opts": _lib.QueryOpts,
               "ebg_mesh_info": _lib.MeshInfo, "ebg_mlp_opts": _lib.MlpOpts, "ebg_prim": _lib.Prim,
               "ebg_batch_opts": _lib.BatchOpts, "ebg_batch_info": _lib.BatchInfo, "ebg_batch_item": _lib.BatchItem}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ebgpu.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


@pytest.mark.parametrize("name", ["dodecahedron.npz", "tetrahedron.npz", "ico3_displaced.npz"])
def test_host_pseudo_normals_match_reference(golden, name):
    from ebgeometry_amd import pseudo_normals
    g = golden(name)
    fn, vn, en = pseudo_normals(g["verts"], g["faces"])
    assert np.abs(fn - g["face_n"]).max() < 1e-14
    assert np.abs(vn - g["vert_n"]).max() < 1e-14
    assert np.abs(en - g["edge_n"]).max() < 1e-14


def test_host_rejects_bad_input():
    from ebgeometry_amd import _lib, pseudo_normals
    v = np.zeros((3, 3), np.float32)
    with pytest.raises(_lib.EbgError):
        pseudo_normals(v, np.array([[0, 1, 5]], np.uint32))  # index out of range
    with pytest.raises(_lib.EbgError):
        pseudo_normals(v, np.array([[0, 1, 2]], np.uint32))  # degenerate (zero area)


def test_mesh_create_validates_without_gpu():
    """Argument validation happens before any device call."""
    from ebgeometry_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 7]], np.uint32)
    st = L.ebg_mesh_create(v.ctypes.data, 3, f.ctypes.data, 1, None, C.byref(h))
    assert st == 1 and not h.value
    v[0, 0] = np.nan
    st = L.ebg_mesh_create(v.ctypes.data, 3, np.array([[0, 1, 2]], np.uint32).ctypes.data, 1, None, C.byref(h))
    assert st == 1


def test_training_schedule_checks_pass_without_a_gpu():
    """Every launch of the training steps ebg_mlp_train enqueues (shapes, split-K chunks and the
    largest index of each buffer access) is checked on the host for the schedules the tests and
    the bench run: the learned-bound test's 400,000 pairs in batches of 2,048 (last batch 640),
    the bench's 4M in 4,096, the gradient tests' single batches of 1,000 and 4,100 rows."""
    from ebgeometry_amd import _lib
    L = _lib.lib()
    for n, b in [(400_000, 2048), (4_000_000, 4096), (1000, 1024), (4100, 4100), (65536, 65536), (100, 32),
                 (8_000_000, 16384), (33, 33)]:
        assert L.ebg_mlp_check_schedule(n, b) == 0, (n, b, L.ebg_last_error())
    assert L.ebg_mlp_check_schedule(1000, 16) != 0  # batch below the kernels' 32-row granularity
