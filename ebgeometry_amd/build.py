"""Build libebgpu.so (HIP kernels for gfx950 + host C++ + the C ABI) in-tree.

    python -m ebgeometry_amd.build            # incremental
    python -m ebgeometry_amd.build --force    # rebuild everything

Objects go to ebgeometry_amd/_build/, the library to ebgeometry_amd/libebgpu.so; the VALU
calibration kernel bench.py uses to price the traversal's instruction issue goes to
ebgeometry_amd/libebgcalib.so.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libebgpu.so")
# measurement aid (bench.py --valu-calibration), not linked into the query library
CALIB_SRC = os.path.join(CSRC, "calib", "valu_calib.hip")
CALIB_LIB = os.path.join(HERE, "libebgcalib.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EBGPU_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
# Device code: gfx950 only, IEEE-rounded FP32 divide/sqrt (the Morton bins depend on it), and
# the packed-FP32 instructions disabled.
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
HOST = ["-pthread"]

SOURCES = [
    ("mesh_build.cpp", HOST),
    ("api.cpp", HOST),
    ("scene_api.cpp", HOST),
    ("csg_api.cpp", HOST),
    ("cloud_api.cpp", HOST),
    ("batch_api.cpp", HOST),
    ("traverse.hip", DEVICE),
    ("scene.hip", DEVICE),
    ("csg.hip", DEVICE),
    ("cloud.hip", DEVICE),
    ("batch.hip", DEVICE),
    ("morton.hip", DEVICE),
    ("morton_build.hip", DEVICE),
    ("gpu_build.hip", DEVICE),
    ("util.hip", DEVICE),
    ("mlp.hip", DEVICE),
    ("noalloc.hip", DEVICE),
]
HEADERS = ["ebgpu_internal.hpp", "kernels.hpp", "mesh_build.hpp", "mlp.hpp", "traverse_dev.hpp", "host_common.hpp",
           "gpu_build.hpp"]


def _mtime(p):
    return os.path.getmtime(p) if os.path.exists(p) else 0.0


def _compile(src, flags, force):
    s = os.path.join(CSRC, src)
    o = os.path.join(BUILD, src + ".o")
    deps = [s] + [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(ROOT, "include", "ebgpu.h")]
    if not force and _mtime(o) >= max(_mtime(d) for d in deps):
        return o, None
    cmd = [HIPCC] + COMMON + flags + ["-c", s, "-o", o]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return o, f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return o, None


def build(force: bool = False, verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        futs = [ex.submit(_compile, s, f, force) for s, f in SOURCES]
        res = [f.result() for f in futs]
    errs = [e for _, e in res if e]
    if errs:
        raise RuntimeError("HIP build failed:\n" + "\n".join(errs))
    objs = [o for o, _ in res]
    if force or _mtime(LIB) < max(_mtime(o) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-pthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        shutil.move(tmp, LIB)
    if force or _mtime(CALIB_LIB) < _mtime(CALIB_SRC):
        cmd = [HIPCC] + COMMON + DEVICE + ["-shared", CALIB_SRC, "-o", CALIB_LIB]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"calibration build failed:\n$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args(argv)
    build(force=a.force, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
