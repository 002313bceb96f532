// GPU Morton build of a mesh BVH: the reference's bottom-up SFC partition
// (TreeBVH::bottomUpSortAndPartition<SFC::Morton>, Source/EBGeometry_BVHImplem.hpp:207-312)
// with triangle bounding boxes as the primitives' volumes (MeshDistanceFunctionsImplem.hpp:
// 133-206):
//   * key of a triangle = Morton code of its box centroid 0.5 (lo + hi) in double, binned over
//     the bounding box of all centroids (SFC::computeBins<double>) -- bit-exact with the
//     reference; stable radix sort on (code, index) (the reference's std::sort leaves equal
//     codes in unspecified order);
//   * a balanced K-ary tree: depth D = floor(log n / log K) (raised so that no leaf holds more
//     than 16 triangles, or lowered by leaf_size), K^D leaves cut along the curve, the first
//     n mod K^D leaves one triangle larger, then K consecutive nodes merged per level.
// Node records are level-ordered: level l starts at (K^l - 1) / (K - 1). Node boxes are
// computed bottom-up, one kernel per level.
#include "kernels.hpp"
#include "mesh_build.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace ebgpu {

namespace {

__global__ __launch_bounds__(256) void tri_centroids_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                            uint64_t nt, double* __restrict__ cen)
{
  const uint64_t f = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= nt) return;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double a = xyz[3 * (size_t)tri[3 * f] + c], b = xyz[3 * (size_t)tri[3 * f + 1] + c],
                 d = xyz[3 * (size_t)tri[3 * f + 2] + c];
    const double lo = fmin(a, fmin(b, d)), hi = fmax(a, fmax(b, d));
    cen[3 * f + c] = 0.5 * (lo + hi);
  }
}

struct Box
{
  float lo[3], hi[3];
};

__device__ __forceinline__ void box_empty(Box& b)
{
  for (int c = 0; c < 3; c++) {
    b.lo[c] = INFINITY;
    b.hi[c] = -INFINITY;
  }
}

// Nodes of the last interior level: child k of node j is leaf j*K + k.
template <int K>
__global__ __launch_bounds__(256) void leaf_level_kernel(const float* __restrict__ xyz, const uint32_t* __restrict__ tri,
                                                         const uint32_t* __restrict__ perm, uint32_t num_nodes,
                                                         uint32_t per, uint32_t rem, Node<K>* __restrict__ level)
{
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= num_nodes) return;
  Node<K>& nd = level[j];
  for (int k = 0; k < K; k++) {
    const uint32_t L     = j * K + k;
    const uint32_t start = L * per + (L < rem ? L : rem);
    const uint32_t cnt   = per + (L < rem ? 1u : 0u);
    Box            b;
    box_empty(b);
    for (uint32_t t = start; t < start + cnt; t++) {
      const uint32_t f = perm[t];
      for (int v = 0; v < 3; v++)
        for (int c = 0; c < 3; c++) {
          const float x = xyz[3 * (size_t)tri[3 * (size_t)f + v] + c];
          b.lo[c]       = fminf(b.lo[c], x);
          b.hi[c]       = fmaxf(b.hi[c], x);
        }
    }
    nd.lox[k] = b.lo[0];
    nd.loy[k] = b.lo[1];
    nd.loz[k] = b.lo[2];
    nd.hix[k] = b.hi[0];
    nd.hiy[k] = b.hi[1];
    nd.hiz[k] = b.hi[2];
    nd.ref[k] = start;
    nd.cnt[k] = cnt;
  }
}

// Nodes of an upper level: child k of node j is node child_first + j*K + k; its box is the
// union of the K boxes that node stores for its own children.
template <int K>
__global__ __launch_bounds__(256) void inner_level_kernel(Node<K>* __restrict__ nodes, uint32_t first,
                                                          uint32_t num_nodes, uint32_t child_first)
{
  const uint32_t j = blockIdx.x * 256 + threadIdx.x;
  if (j >= num_nodes) return;
  Node<K>& nd = nodes[first + j];
  for (int k = 0; k < K; k++) {
    const uint32_t c  = child_first + j * K + k;
    const Node<K>& ch = nodes[c];
    Box            b;
    box_empty(b);
    for (int i = 0; i < K; i++) {
      b.lo[0] = fminf(b.lo[0], ch.lox[i]);
      b.lo[1] = fminf(b.lo[1], ch.loy[i]);
      b.lo[2] = fminf(b.lo[2], ch.loz[i]);
      b.hi[0] = fmaxf(b.hi[0], ch.hix[i]);
      b.hi[1] = fmaxf(b.hi[1], ch.hiy[i]);
      b.hi[2] = fmaxf(b.hi[2], ch.hiz[i]);
    }
    nd.lox[k] = b.lo[0];
    nd.loy[k] = b.lo[1];
    nd.loz[k] = b.lo[2];
    nd.hix[k] = b.hi[0];
    nd.hiy[k] = b.hi[1];
    nd.hiz[k] = b.hi[2];
    nd.ref[k] = c;
    nd.cnt[k] = 0;
  }
}

struct Dev
{
  void* p = nullptr;
  ~Dev()
  {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(std::size_t b) { return hipMalloc(&p, std::max<std::size_t>(b, 256)); }
  template <class T>
  T* as() const
  {
    return static_cast<T*>(p);
  }
};

#define EBG_TRY(x)                                                                                            \
  do {                                                                                                        \
    const hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) {                                                                                   \
      err = std::string(#x) + ": " + hipGetErrorString(e_);                                                  \
      return e_;                                                                                              \
    }                                                                                                         \
  } while (0)

template <int K>
hipError_t build_levels(const float* d_xyz, const uint32_t* d_tri, const uint32_t* d_perm, uint64_t nt, int depth,
                        std::vector<uint8_t>& nodes_out, std::size_t& num_nodes, std::string& err)
{
  // level offsets
  std::vector<uint64_t> off(depth + 1), cnt(depth + 1);
  uint64_t              total = 0, w = 1;
  for (int l = 0; l < depth; l++) {
    off[l] = total;
    cnt[l] = w;
    total += w;
    w *= K;
  }
  const uint64_t leaves = w; // K^depth
  const uint32_t per    = uint32_t(nt / leaves);
  const uint32_t rem    = uint32_t(nt % leaves);
  Dev            d_nodes;
  EBG_TRY(d_nodes.alloc(total * sizeof(Node<K>)));
  Node<K>* nd = d_nodes.as<Node<K>>();
  const int l_last = depth - 1;
  hipLaunchKernelGGL(leaf_level_kernel<K>, dim3((uint32_t)((cnt[l_last] + 255) / 256)), dim3(256), 0, nullptr, d_xyz,
                     d_tri, d_perm, (uint32_t)cnt[l_last], per, rem, nd + off[l_last]);
  EBG_TRY(hipGetLastError());
  for (int l = l_last - 1; l >= 0; l--) {
    hipLaunchKernelGGL(inner_level_kernel<K>, dim3((uint32_t)((cnt[l] + 255) / 256)), dim3(256), 0, nullptr, nd,
                       (uint32_t)off[l], (uint32_t)cnt[l], (uint32_t)off[l + 1]);
    EBG_TRY(hipGetLastError());
  }
  nodes_out.resize(total * sizeof(Node<K>));
  EBG_TRY(hipMemcpy(nodes_out.data(), nd, nodes_out.size(), hipMemcpyDeviceToHost));
  num_nodes = total;
  return hipSuccess;
}

} // namespace

int morton_depth(uint64_t nt, int K, int leaf_size)
{
  // the reference's depth (Source/EBGeometry_BVHImplem.hpp:242-245) for leaf_size 1
  const uint64_t base  = std::max<uint64_t>(1, nt / (uint64_t)std::max(1, leaf_size));
  int            depth = (int)std::floor(std::log((double)base) / std::log((double)K));
  if (depth < 0) depth = 0;
  auto leaves = [&](int d) {
    uint64_t w = 1;
    for (int i = 0; i < d; i++) w *= (uint64_t)K;
    return w;
  };
  while ((nt + leaves(depth) - 1) / leaves(depth) > (uint64_t)kMaxLeaf) depth++;
  while (depth > 0 && leaves(depth) > nt) depth--;
  return depth;
}

bool build_morton_mesh_gpu(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt, int K, int leaf_size,
                           const PseudoNormals& pn, PackedMesh& out, std::string& err)
{
  if (K != 4 && K != 8) {
    err = "branching must be 4 or 8";
    return false;
  }
  const int depth = morton_depth(nt, K, leaf_size);
  Dev       d_xyz, d_tri, d_cen, d_box, d_codes, d_sorted, d_perm, d_scratch;
  auto      run   = [&]() -> hipError_t {
    EBG_TRY(d_xyz.alloc(12 * nv));
    EBG_TRY(d_tri.alloc(12 * nt));
    EBG_TRY(d_cen.alloc(24 * nt));
    EBG_TRY(d_box.alloc(sizeof(double) * 6 * (kBboxParts + 1)));
    EBG_TRY(d_codes.alloc(8 * nt));
    EBG_TRY(d_sorted.alloc(8 * nt));
    EBG_TRY(d_perm.alloc(4 * nt));
    EBG_TRY(d_scratch.alloc(radix_sort_scratch_bytes(nt)));
    EBG_TRY(hipMemcpy(d_xyz.p, xyz, 12 * nv, hipMemcpyHostToDevice));
    EBG_TRY(hipMemcpy(d_tri.p, tri, 12 * nt, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(tri_centroids_kernel, dim3((uint32_t)((nt + 255) / 256)), dim3(256), 0, nullptr,
                       d_xyz.as<float>(), d_tri.as<uint32_t>(), (uint64_t)nt, d_cen.as<double>());
    EBG_TRY(hipGetLastError());
    double* box = d_box.as<double>();
    EBG_TRY(launch_bbox_f64(d_cen.as<double>(), nt, box, box + 8, nullptr));
    EBG_TRY(launch_codes_f64(d_cen.as<double>(), nt, box, d_codes.as<uint64_t>(), nullptr));
    EBG_TRY(radix_sort_pairs(d_codes.as<uint64_t>(), nullptr, d_sorted.as<uint64_t>(), d_perm.as<uint32_t>(), nt, 63,
                             d_scratch.p, nullptr));
    std::vector<uint32_t> perm(nt);
    if (depth == 0) {
      // fewer than K triangles: one root with a single leaf
      Node<4> n4{};
      Node<8> n8{};
      float   lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      for (std::size_t f = 0; f < nt; f++)
        for (int v = 0; v < 3; v++)
          for (int c = 0; c < 3; c++) {
            lo[c] = std::min(lo[c], xyz[3 * tri[3 * f + v] + c]);
            hi[c] = std::max(hi[c], xyz[3 * tri[3 * f + v] + c]);
          }
      auto fill = [&](auto& n, int KK) {
        for (int k = 0; k < KK; k++) {
          n.cnt[k] = kInvalid;
          n.ref[k] = 0;
          n.lox[k] = n.loy[k] = n.loz[k] = INFINITY;
          n.hix[k] = n.hiy[k] = n.hiz[k] = -INFINITY;
        }
        n.lox[0] = lo[0], n.loy[0] = lo[1], n.loz[0] = lo[2];
        n.hix[0] = hi[0], n.hiy[0] = hi[1], n.hiz[0] = hi[2];
        n.ref[0] = 0;
        n.cnt[0] = uint32_t(nt);
      };
      if (K == 4) {
        fill(n4, 4);
        out.nodes.resize(sizeof(n4));
        std::memcpy(out.nodes.data(), &n4, sizeof(n4));
      }
      else {
        fill(n8, 8);
        out.nodes.resize(sizeof(n8));
        std::memcpy(out.nodes.data(), &n8, sizeof(n8));
      }
      out.num_nodes  = 1;
      out.num_leaves = 1;
    }
    else {
      std::size_t num_nodes = 0;
      const hipError_t e = K == 4 ? build_levels<4>(d_xyz.as<float>(), d_tri.as<uint32_t>(), d_perm.as<uint32_t>(), nt,
                                                    depth, out.nodes, num_nodes, err)
                                  : build_levels<8>(d_xyz.as<float>(), d_tri.as<uint32_t>(), d_perm.as<uint32_t>(), nt,
                                                    depth, out.nodes, num_nodes, err);
      if (e != hipSuccess) return e;
      out.num_nodes = num_nodes;
      uint64_t w    = 1;
      for (int i = 0; i < depth; i++) w *= (uint64_t)K;
      out.num_leaves = w;
    }
    EBG_TRY(hipMemcpy(perm.data(), d_perm.p, 4 * nt, hipMemcpyDeviceToHost));
    out.K         = K;
    out.leaf_size = leaf_size;
    // root box = union of the root's child boxes
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    const float* r = reinterpret_cast<const float*>(out.nodes.data());
    for (int k = 0; k < K; k++) {
      const uint32_t c = reinterpret_cast<const uint32_t*>(r)[7 * K + k];
      if (c == kInvalid) continue;
      for (int a = 0; a < 3; a++) {
        lo[a] = std::min(lo[a], r[a * K + k]);
        hi[a] = std::max(hi[a], r[(3 + a) * K + k]);
      }
    }
    for (int a = 0; a < 3; a++) {
      out.lo[a] = lo[a];
      out.hi[a] = hi[a];
    }
    const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    out.diag        = std::sqrt(dx * dx + dy * dy + dz * dz);
    pack_triangles(xyz, tri, perm, pn, nv, out);
    out.stack_need = compute_stack_need(out);
    return hipSuccess;
  };
  if (run() != hipSuccess) return false;
  return true;
}

hipError_t launch_tri_centroids(const float* d_xyz, const uint32_t* d_tri, uint64_t nt, double* d_cen, hipStream_t s)
{
  if (nt == 0) return hipSuccess;
  hipLaunchKernelGGL(tri_centroids_kernel, dim3((uint32_t)((nt + 255) / 256)), dim3(256), 0, s, d_xyz, d_tri, nt, d_cen);
  return hipGetLastError();
}

} // namespace ebgpu
