"""The host-side mesh preparation (ebgeometry_amd/csrc/mesh_build.cpp: pseudo-normals, polygon
clipping, binned SAH, collapse to K-wide nodes, packing, stack bound, the box BVH of point
clouds) built with AddressSanitizer and UndefinedBehaviorSanitizer (g++, host code only) and run
over meshes of every size class, every branching factor and leaf size, and the degenerate inputs
the library must refuse. Any out-of-bounds access, use of freed memory or undefined operation
aborts the program."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

CSRC = os.path.join(ROOT, "ebgeometry_amd", "csrc")

DRIVER = r"""
#include "mesh_build.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

using namespace ebgpu;

#define CHECK(c)                                                                                  \
  do {                                                                                            \
    if (!(c)) {                                                                                   \
      std::printf("check failed at line %d: %s\n", __LINE__, #c);                                 \
      std::exit(1);                                                                               \
    }                                                                                             \
  } while (0)

// unit icosphere subdivided `level` times
static void icosphere(int level, std::vector<float>& v, std::vector<uint32_t>& f)
{
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<double> p = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                           0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
  std::vector<uint32_t> tri = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4, 11, 10, 2, 10, 7, 6,
                               7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8, 3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7,
                               9, 8, 1};
  for (int l = 0; l < level; l++) {
    std::vector<uint32_t> nt;
    auto mid = [&](uint32_t a, uint32_t b) {
      p.push_back(0.5 * (p[3 * a] + p[3 * b]));
      p.push_back(0.5 * (p[3 * a + 1] + p[3 * b + 1]));
      p.push_back(0.5 * (p[3 * a + 2] + p[3 * b + 2]));
      return uint32_t(p.size() / 3 - 1);
    };
    for (size_t i = 0; i < tri.size(); i += 3) {
      // unshared midpoints: a valid (if not watertight) triangle soup for the builders
      const uint32_t a = tri[i], b = tri[i + 1], c = tri[i + 2];
      const uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      const uint32_t q[12] = {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca};
      nt.insert(nt.end(), q, q + 12);
    }
    tri.swap(nt);
  }
  v.resize(p.size());
  for (size_t i = 0; i < p.size(); i += 3) {
    const double r = std::sqrt(p[i] * p[i] + p[i + 1] * p[i + 1] + p[i + 2] * p[i + 2]);
    for (int c = 0; c < 3; c++) v[i + c] = float(p[i + c] / r);
  }
  f = tri;
}

static void build_all(const std::vector<float>& v, const std::vector<uint32_t>& f)
{
  const size_t         nv = v.size() / 3, nt = f.size() / 3;
  PseudoNormals        pn;
  std::string          err;
  CHECK(compute_pseudo_normals(v.data(), nv, f.data(), nt, pn, err));
  for (int K : {4, 8})
    for (int leaf : {1, 3, 8, 16}) {
      PackedMesh pm;
      CHECK(build_sah_mesh(v.data(), nv, f.data(), nt, K, leaf, pn, pm, err));
      CHECK(pm.tris.size() == nt && pm.order.size() == nt);
      CHECK(pm.stack_need >= 1 && pm.num_nodes >= 1);
      CHECK(compute_stack_need(pm) == pm.stack_need);
      std::vector<char> seen(nt, 0);
      for (uint32_t o : pm.order) {
        CHECK(o < nt && !seen[o]);
        seen[o] = 1;
      }
    }
}

int main()
{
  // triangle meshes from 1 to 81,920 faces
  for (int level : {0, 1, 3, 6}) {
    std::vector<float>    v;
    std::vector<uint32_t> f;
    icosphere(level, v, f);
    build_all(v, f);
    if (level == 0) {
      build_all(v, std::vector<uint32_t>(f.begin(), f.begin() + 3));
      build_all(v, std::vector<uint32_t>(f.begin(), f.begin() + 6));
    }
  }
  // refused: a zero-area triangle and a repeated vertex
  {
    std::vector<float>    v = {0, 0, 0, 1, 0, 0, 2, 0, 0, 0, 1, 0};
    std::vector<uint32_t> f = {0, 1, 2};
    PseudoNormals         pn;
    std::string           err;
    CHECK(!compute_pseudo_normals(v.data(), 4, f.data(), 1, pn, err));
    f = {0, 1, 1};
    CHECK(!compute_pseudo_normals(v.data(), 4, f.data(), 1, pn, err));
  }
  // polygon meshes: the unit cube in quads, an L-shaped prism (non-convex hexagon caps)
  {
    std::vector<float>    v;
    for (int i = 0; i < 8; i++) v.insert(v.end(), {float(i >> 2 & 1), float(i >> 1 & 1), float(i & 1)});
    std::vector<uint32_t> off = {0, 4, 8, 12, 16, 20, 24};
    std::vector<uint32_t> idx = {0, 1, 3, 2, 4, 6, 7, 5, 0, 4, 5, 1, 2, 3, 7, 6, 0, 2, 6, 4, 1, 5, 7, 3};
    std::vector<uint32_t> tri, tf;
    PseudoNormals         pn;
    std::string           err;
    CHECK(polygon_mesh(v.data(), 8, off.data(), idx.data(), 6, tri, tf, pn, err));
    CHECK(tf.size() == 12 && tri.size() == 36 && pn.edge.size() == 9 * 12);
    PackedMesh pm;
    CHECK(build_sah_mesh(v.data(), 8, tri.data(), tf.size(), 4, 2, pn, pm, err));
    const float L[12] = {0, 0, 2, 0, 2, 0.7f, 0.7f, 0.7f, 0.7f, 2, 0, 2};
    std::vector<float> w;
    for (int z = 0; z < 2; z++)
      for (int k = 0; k < 6; k++) w.insert(w.end(), {L[2 * k], L[2 * k + 1], float(z)});
    std::vector<uint32_t> o2 = {0, 6, 12}, i2 = {5, 4, 3, 2, 1, 0, 6, 7, 8, 9, 10, 11};
    for (uint32_t k = 0; k < 6; k++) {
      i2.insert(i2.end(), {k, (k + 1) % 6, 6 + (k + 1) % 6, 6 + k});
      o2.push_back(uint32_t(i2.size()));
    }
    CHECK(polygon_mesh(w.data(), 12, o2.data(), i2.data(), 8, tri, tf, pn, err));
    CHECK(tf.size() == 2 * 4 + 6 * 2);
    // refused: a bow tie, a non-planar quad, a face of two corners, a repeated corner
    std::vector<float>    b = {0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0, 0, 1, 0.3f};
    std::vector<uint32_t> bo = {0, 4}, bi = {0, 1, 2, 3};
    CHECK(!polygon_mesh(b.data(), 5, bo.data(), bi.data(), 1, tri, tf, pn, err));
    bi = {0, 1, 3, 4};
    CHECK(!polygon_mesh(b.data(), 5, bo.data(), bi.data(), 1, tri, tf, pn, err));
    bo = {0, 2};
    CHECK(!polygon_mesh(b.data(), 5, bo.data(), bi.data(), 1, tri, tf, pn, err));
    bo = {0, 4}, bi = {0, 1, 3, 1};
    CHECK(!polygon_mesh(b.data(), 5, bo.data(), bi.data(), 1, tri, tf, pn, err));
  }
  // point-cloud trees over random boxes (equal points included)
  {
    std::mt19937       g(7);
    std::normal_distribution<float> nd;
    for (size_t n : {1ul, 2ul, 17ul, 5000ul, 100000ul}) {
      std::vector<float> p(3 * n);
      for (auto& x : p) x = nd(g);
      if (n > 10) std::copy(p.begin(), p.begin() + 30, p.end() - 30);
      for (int K : {4, 8})
        for (int leaf : {1, 8, 16}) {
          WideBvh     wb;
          std::string err;
          CHECK(build_wide_bvh(p.data(), p.data(), n, K, leaf, wb, err));
          CHECK(wb.order.size() == n && wb.num_nodes >= 1);
        }
    }
  }
  std::printf("host build checks passed\n");
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_host_build_under_asan_and_ubsan(tmp_path):
    src = tmp_path / "driver.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-pthread", "-I" + CSRC, str(src), os.path.join(CSRC, "mesh_build.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "host build checks passed" in run.stdout
