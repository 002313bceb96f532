"""The C++ adapters (include/ebgpu_sdf.hpp) compile against the reference's own headers and link
with libebgpu.so: GpuTriMeshSDF<T> is a SignedDistanceFunction<T> and GpuUnionIF<T> an
ImplicitFunction<T> (Source/EBGeometry_SignedDistanceFunction.hpp:54-62,
Source/EBGeometry_ImplicitFunction.hpp:51-52), held as the reference's callers hold theirs.
Without a GPU the program exercises the host-only entry point and the adapters' error path."""
import os
import subprocess

import pytest

from conftest import ROOT

EBG_HOME = os.environ.get("EBGEOMETRY_HOME", "/root/reference")

PROGRAM = r"""
#include <EBGeometry.hpp>
#include "ebgpu_sdf.hpp"
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

int main()
{
  using V = EBGeometry::Vec3T<double>;
  // a tetrahedron, outward oriented
  std::vector<V> v = {V(0, 0, 0), V(1, 0, 0), V(0, 1, 0), V(0, 0, 1)};
  std::vector<std::array<uint32_t, 3>> t = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
  // host-only entry point: the pseudo-normals of the first face point away from the solid
  std::vector<float> xyz;
  for (const auto& p : v) for (int c = 0; c < 3; c++) xyz.push_back(float(p[c]));
  std::vector<uint32_t> tri;
  for (const auto& f : t) for (int c = 0; c < 3; c++) tri.push_back(f[c]);
  std::vector<double> fn(12), vn(36), en(36);
  if (ebg_pseudo_normals_host(xyz.data(), 4, tri.data(), 4, fn.data(), vn.data(), en.data()) != EBG_OK) return 2;
  if (!(fn[2] < -0.99)) return 3;
  try {
    std::shared_ptr<EBGeometry::ImplicitFunction<double>> f = std::make_shared<ebgpu::GpuTriMeshSDF<double>>(v, t);
    std::shared_ptr<EBGeometry::ImplicitFunction<float>> g =
      std::make_shared<ebgpu::GpuTriMeshSDF<float>>(std::vector<EBGeometry::Vec3T<float>>{
        EBGeometry::Vec3T<float>(0, 0, 0), EBGeometry::Vec3T<float>(1, 0, 0), EBGeometry::Vec3T<float>(0, 1, 0),
        EBGeometry::Vec3T<float>(0, 0, 1)}, t);
    // the unit cube with quad faces, as MeshSDF<double, Meta, K> takes it
    std::vector<V> cv;
    for (int i = 0; i < 8; i++) cv.push_back(V(i >> 2 & 1, i >> 1 & 1, i & 1));
    std::vector<std::vector<size_t>> quads = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1},
                                              {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
    std::shared_ptr<EBGeometry::SignedDistanceFunction<double>> c =
      std::make_shared<ebgpu::GpuMeshSDF<double>>(cv, quads);
    const double d = f->value(V(2, 2, 2)), dc = c->signedDistance(V(0.5, 0.5, 0.25));
    std::printf("device %.6f %.6f %.6f\n", d, double(g->value(EBGeometry::Vec3T<float>(2, 2, 2))), dc);
    return std::isfinite(d) && std::abs(dc + 0.25) < 1e-6 ? 0 : 4;
  }
  catch (const std::runtime_error& e) {
    std::printf("no device: %s\n", e.what());
    return 0;
  }
}
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(EBG_HOME, "EBGeometry.hpp")), reason="reference headers absent")
def test_adapters_compile_against_reference_headers(tmp_path):
    lib_dir = os.path.join(ROOT, "ebgeometry_amd")
    assert os.path.exists(os.path.join(lib_dir, "libebgpu.so")), "build the library first"
    src = tmp_path / "adapter.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "adapter"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", EBG_HOME, "-I", os.path.join(ROOT, "include"), str(src),
                    "-o", str(exe), "-L", lib_dir, "-lebgpu", "-Wl,-rpath," + lib_dir], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.startswith(("device ", "no device: ebgpu: ")), r.stdout
