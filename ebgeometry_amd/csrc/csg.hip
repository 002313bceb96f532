// CSG programs on gfx950: the reference's implicit-function algebra evaluated per query.
//
// A program is the post-order instruction list of an expression tree (CsgOpcode): analytic
// leaves (AnalyticDistanceFunctions.hpp), mesh leaves (TriMeshSDF), point transforms
// (Translate / Rotate / Scale / Elongate / Reflect, Transform.hpp; FiniteRepetition, CSG.hpp:805),
// value transforms (Complement / Offset / Annular / the factor of Scale), folds for the n-ary
// operators (Union / Intersection / SmoothUnion / SmoothIntersection / Difference, CSG.hpp:293-803;
// Blur and Mollify are weighted sums) and BVH unions over primitive subprograms (BVHUnionIF,
// BVHSmoothUnionIF, CSGImplem.hpp:466-608).
//
// One query per lane; every lane runs the same instruction sequence, so control flow and the
// stack depths are wave-uniform: instructions are read with scalar loads, the value stack and
// the saved points live in LDS (one column per lane, no cross-lane traffic, no barriers). A BVH
// union is walked by the whole wavefront as the scene union is (wave-uniform VGPR-lane stack,
// node boxes tested per lane against max(0, v)^2, v the running minimum -- the second smallest
// for the smooth union -- as pruneTraverse does, BVHImplem.hpp:1098-1606); a primitive some
// lane needs is evaluated for the wavefront by running its subprogram. A mesh leaf is walked by
// the wavefront with the mesh traversal.
//
// FP32 pass for every query; the queries with |v| below the recheck threshold are evaluated
// again in FP64 (FP64 parameters, FP64 triangle math).
#include "kernels.hpp"
#include "traverse_dev.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace ebgpu {

namespace {

using namespace dev;

// Per-workgroup LDS of W lanes: FP32 40 KB at W = 256, FP64 40 KB at W = 128.
template <class R, uint32_t W>
struct CsgLds
{
  R v[kCsgMaxValues][W];    // value stack
  R p[kCsgMaxPoints][3][W]; // saved points
};

// One lane's interpreter state; its stack entries are vcol[i * W], pcol[(3 d + c) * W].
template <class R>
struct CsgLane
{
  R        x, y, z;    // current point
  int      vtop, ptop; // wave-uniform depths
  R*       vcol;
  R*       pcol;
  uint32_t w;          // lanes per workgroup (the column stride)
};

template <class R>
struct CsgParams
{
  uint32_t op, n, a, b;
  R        p[12];
};

template <class R>
__device__ __forceinline__ CsgParams<R> fetch(const CsgDev& g, uint32_t i)
{
  CsgParams<R> q;
  if constexpr (sizeof(R) == 4) {
    const CsgInstr in = sload_record(g.ops + i);
    q.op = in.op, q.n = in.n, q.a = in.a, q.b = in.b;
#pragma unroll
    for (int k = 0; k < 12; k++) q.p[k] = in.p[k];
  }
  else {
    const CsgInstr64 in = sload_record(g.ops64 + i);
    q.op = in.op, q.n = in.n, q.a = in.a, q.b = in.b;
#pragma unroll
    for (int k = 0; k < 12; k++) q.p[k] = in.p[k];
  }
  return q;
}

template <class R>
__device__ __forceinline__ R clampv(R x, R lo, R hi)
{
  return x < lo ? lo : (hi < x ? hi : x); // std::clamp
}

template <class R>
__device__ __forceinline__ R len3(R x, R y, R z)
{
  return sqrt(x * x + y * y + z * z);
}

// Analytic leaves (AnalyticDistanceFunctions.hpp; line numbers of signedDistance).
template <class R>
__device__ __forceinline__ R leaf_value(const CsgParams<R>& q, R x, R y, R z)
{
  const R* p = q.p;
  switch (q.op) {
  case kCsgSphere: // :240-247
    return len3(x - p[0], y - p[1], z - p[2]) - p[3];
  case kCsgBox: { // :377-399
    const R dx = max(p[0] - x, x - p[3]), dy = max(p[1] - y, y - p[4]), dz = max(p[2] - z, z - p[5]);
    return min(R(0), max(dx, max(dy, dz))) + len3(max(R(0), dx), max(R(0), dy), max(R(0), dz));
  }
  case kCsgPlane: // :107-116
    return (x - p[0]) * p[3] + (y - p[1]) * p[4] + (z - p[2]) * p[5];
  case kCsgTorus: { // :551-562
    const R ux = x - p[0], uy = y - p[1], uz = z - p[2];
    const R rho = sqrt(ux * ux + uy * uy) - p[3];
    return sqrt(rho * rho + uz * uz) - p[4];
  }
  case kCsgCylinder: { // :694-734
    if (!(p[6] > R(0) && p[7] > R(0))) return R(INFINITY);
    const R ux = x - p[0], uy = y - p[1], uz = z - p[2];
    const R para = ux * p[3] + uy * p[4] + uz * p[5];
    const R w    = len3(ux - para * p[3], uy - para * p[4], uz - para * p[5]) - p[7];
    const R h    = fabs(para) - R(0.5) * p[6];
    if (w <= R(0) && h <= R(0)) return fabs(w) < fabs(h) ? w : h;
    if (w <= R(0) && h > R(0)) return h;
    if (w > R(0) && h < R(0)) return w;
    return sqrt(w * w + h * h);
  }
  case kCsgInfCylinder: { // :844-856
    R ux = x - p[0], uy = y - p[1], uz = z - p[2];
    if (q.n == 0) ux = R(0);
    else if (q.n == 1) uy = R(0);
    else if (q.n == 2) uz = R(0);
    return len3(ux, uy, uz) - p[3];
  }
  case kCsgCapsule: { // :970-984
    const R v1x = x - p[0], v1y = y - p[1], v1z = z - p[2];
    const R v2x = p[3] - p[0], v2y = p[4] - p[1], v2z = p[5] - p[2];
    const R h   = clampv((v1x * v2x + v1y * v2y + v1z * v2z) / (v2x * v2x + v2y * v2y + v2z * v2z), R(0), R(1));
    return len3(v1x - h * v2x, v1y - h * v2y, v1z - h * v2z) - p[6];
  }
  case kCsgInfCone: { // :1085-1099
    const R dx = x - p[0], dy = y - p[1], dz = z - p[2];
    const R qx = sqrt(dx * dx + dy * dy), qy = -dz;
    const R t  = max(qx * p[3] + qy * p[4], R(0));
    const R ex = qx - p[3] * t, ey = qy - p[4] * t;
    const R d1 = sqrt(ex * ex + ey * ey);
    return d1 * ((qx * p[4] - qy * p[3] < R(0)) ? R(-1) : R(1));
  }
  case kCsgCone: { // :1195-1222
    const R dx = x - p[0], dy = y - p[1], dz = z - p[2];
    const R wx = sqrt(dx * dx + dy * dy), wy = dz;
    const R qx = p[5] * (p[3] / p[4]), qy = p[5] * R(-1);
    const R ta = clampv((wx * qx + wy * qy) / (qx * qx + qy * qy), R(0), R(1));
    const R ax = wx - ta * qx, ay = wy - ta * qy;
    const R bx = wx - qx * clampv(wx / qx, R(0), R(1)), by = wy - qy;
    const R k  = R((qy > R(0)) - (qy < R(0)));
    const R d  = min(ax * ax + ay * ay, bx * bx + by * by);
    const R s  = max(k * (wx * qy - wy * qx), k * (wy - qy));
    return sqrt(d) * R((s > R(0)) - (s < R(0)));
  }
  case kCsgRoundedBox: // :1320-1327
    return len3(x - clampv(x, -p[0], p[0]), y - clampv(y, -p[1], p[1]), z - clampv(z, -p[2], p[2])) - p[3];
  case kCsgRoundedCylinder: { // :1721-1733
    const R xz  = sqrt(z * z + x * x);
    const R d1x = xz - p[0], d1y = fabs(y) - p[2];
    const R d2x = max(d1x, R(0)), d2y = max(d1y, R(0));
    return min(max(d1x, d1y), R(0)) + sqrt(d2x * d2x + d2y * d2y) - p[1];
  }
  default:
    return R(NAN);
  }
}

template <class R>
__device__ __forceinline__ R smooth_min(R a, R b, R s, uint32_t kind)
{
  if (kind == 1u) return -log(exp(-a / s) + exp(-b / s)) * s; // ExpMin, CSG.hpp:236-245
  const R h = max(s - fabs(a - b), R(0)) / s;                   // SmoothMin, CSG.hpp:256-263
  return min(a, b) - R(0.25) * h * h * s;
}

template <class R>
__device__ __forceinline__ R smooth_max(R a, R b, R s)
{
  const R h = max(s - fabs(a - b), R(0)) / s; // SmoothMax, CSG.hpp:276-283
  return max(a, b) + R(0.25) * h * h * s;
}

template <class R>
__device__ __forceinline__ R& vat(CsgLane<R>& L, int i)
{
  return L.vcol[i * L.w];
}

template <class R>
__device__ __forceinline__ R& pat(CsgLane<R>& L, int d, int c)
{
  return L.pcol[(3 * d + c) * L.w];
}

// One instruction other than a BVH union. `need`: this lane wants the value (mesh walks leave
// the other lanes out).
template <class R, int KM, int NB>
__device__ __forceinline__ void csg_step(const CsgDev& g, const CsgParams<R>& q, CsgLane<R>& L, bool need)
{
  const uint32_t op = q.op;
  if (op <= kCsgRoundedCylinder) {
    vat(L, L.vtop++) = leaf_value<R>(q, L.x, L.y, L.z);
    return;
  }
  if (op == kCsgMesh) {
    const MeshDev md   = sload_record(g.meshes + q.n);
    R             best = R(INFINITY);
    uint32_t      btri = kInvalid, bfeat = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    traverse<KM, NB, R, false, true>(md, L.x, L.y, L.z, need ? R(INFINITY) : R(-1), best, btri, bfeat, c0, c1, c2, c3);
    R val = R(INFINITY);
    if (need && btri != kInvalid) {
      if constexpr (sizeof(R) == 4) val = resolve32(md, btri, bfeat, best, L.x, L.y, L.z);
      else val = resolve64(md, btri, bfeat, best, L.x, L.y, L.z);
    }
    vat(L, L.vtop++) = val;
    return;
  }
  if (op >= kCsgTranslate && op <= kCsgRepeat) {
    const int d = L.ptop++;
    pat(L, d, 0) = L.x;
    pat(L, d, 1) = L.y;
    pat(L, d, 2) = L.z;
    const R* p        = q.p;
    const R  x = L.x, y = L.y, z = L.z;
    switch (op) {
    case kCsgTranslate: // TranslateIF, TransformImplem.hpp:154-162
      L.x = x - p[0], L.y = y - p[1], L.z = z - p[2];
      break;
    case kCsgRotate: // RotateIF, TransformImplem.hpp:185-226
      if (q.n == 0) L.y = y * p[0] + z * p[1], L.z = -y * p[1] + z * p[0];
      else if (q.n == 1) L.x = x * p[0] - z * p[1], L.z = x * p[1] + z * p[0];
      else if (q.n == 2) L.x = x * p[0] + y * p[1], L.y = -x * p[1] + y * p[0];
      break;
    case kCsgScalePoint: // ScaleIF, TransformImplem.hpp:262-270
      L.x = x / p[0], L.y = y / p[0], L.z = z / p[0];
      break;
    case kCsgElongate: // ElongateIF, TransformImplem.hpp:427-435
      L.x = x - clampv(x, -p[0], p[0]), L.y = y - clampv(y, -p[1], p[1]), L.z = z - clampv(z, -p[2], p[2]);
      break;
    case kCsgReflect: // ReflectIF, TransformImplem.hpp:454-462
      L.x = x * p[0], L.y = y * p[1], L.z = z * p[2];
      break;
    default: // kCsgRepeat, FiniteRepetitionIF, CSGImplem.hpp:815-826
      L.x = x - p[0] * round(clampv(x / p[0], -p[3], p[6]));
      L.y = y - p[1] * round(clampv(y / p[1], -p[4], p[7]));
      L.z = z - p[2] * round(clampv(z / p[2], -p[5], p[8]));
      break;
    }
    return;
  }
  if (op == kCsgPopPoint) {
    const int d = --L.ptop;
    L.x         = pat(L, d, 0);
    L.y         = pat(L, d, 1);
    L.z         = pat(L, d, 2);
    return;
  }
  if (op == kCsgPushConst) {
    vat(L, L.vtop++) = q.p[0];
    return;
  }
  R& u = vat(L, L.vtop - 1);
  switch (op) {
  case kCsgScaleValue: u = u * q.p[0]; return;         // ScaleIF
  case kCsgNegate: u = -u; return;                      // ComplementIF, TransformImplem.hpp:129-137
  case kCsgOffset: u = u - q.p[0]; return;              // OffsetIF, :239-247
  case kCsgAnnular: u = fabs(u) - q.p[0]; return;       // AnnularIF, :284-292
  case kCsgSmoothMaxEnd: {
    const R a = vat(L, L.vtop - 2);
    L.vtop--;
    vat(L, L.vtop - 1) = smooth_max<R>(a, u, q.p[0]);
    return;
  }
  case kCsgSmoothMinEnd: {
    const R a = vat(L, L.vtop - 2);
    L.vtop--;
    vat(L, L.vtop - 1) = smooth_min<R>(a, u, q.p[0], q.n);
    return;
  }
  case kCsgDifference: { // DifferenceIF, CSGImplem.hpp:735-741
    const R a = vat(L, L.vtop - 2);
    L.vtop--;
    vat(L, L.vtop - 1) = max(a, -u);
    return;
  }
  default:
    break;
  }
  // folds: pop v into the accumulator(s) below
  const R v = u;
  L.vtop--;
  if (op == kCsgMinAcc) { R& w = vat(L, L.vtop - 1); w = min(w, v); }      // UnionIF, CSGImplem.hpp:355-366
  else if (op == kCsgMaxAcc) { R& w = vat(L, L.vtop - 1); w = max(w, v); } // IntersectionIF, :624-635
  else if (op == kCsgAddAcc) { R& w = vat(L, L.vtop - 1); w = w + v; }
  else if (op == kCsgTwoMinAcc) { // SmoothUnionIF, :390-414
    R& a = vat(L, L.vtop - 2);
    R& b = vat(L, L.vtop - 1);
    if (v < a) { b = a; a = v; }
    else if (v < b) b = v;
  }
  else if (op == kCsgTwoMaxAcc) { // SmoothIntersectionIF, :678-702
    R& a = vat(L, L.vtop - 2);
    R& b = vat(L, L.vtop - 1);
    if (v > a) { b = a; a = v; }
    else if (v > b) b = v;
  }
}

// BVH union over primitive subprograms: pushes min over the primitives (or the smooth min of
// the two smallest).
template <class R, int KM, int NB>
__device__ __forceinline__ void csg_bvh(const CsgDev& g, const CsgParams<R>& q, CsgLane<R>& L, bool active)
{
  const CsgBvhDev B      = sload_record(g.bvhs + q.a);
  const bool      smooth = q.op == kCsgBvhSmoothUnion;
  const Node<4>* __restrict__ nodes = static_cast<const Node<4>*>(B.nodes);
  R               a = R(INFINITY), b = R(INFINITY);
  R               bound = active ? R(INFINITY) : R(-1);
  LaneStack<NB>   st;
  st.init();
  const int lane = __lane_id();
  int       top  = 0;
  st.put(top++, Entry{0u, 0u, B.root_lo[0], B.root_lo[1], B.root_lo[2], B.root_hi[0], B.root_hi[1], B.root_hi[2]});
  while (top > 0) {
    const Entry e    = st.get(--top);
    const bool  need = box_d2<R>(L.x, L.y, L.z, e.lx, e.ly, e.lz, e.hx, e.hy, e.hz) <= bound;
    if (ballot(need) == 0) continue;
    if (e.cnt == 0u) {
      push_children<4, NB, R, true>(sload_record(nodes + uni(e.ref)), L.x, L.y, L.z, bound, st, top, 0);
      if (top > 64 * NB) {
        if (lane == 0 && g.status) atomicOr(g.status, kStatusStackOverflow);
        top = 64 * NB;
      }
      continue;
    }
    const uint32_t first = uni(e.ref), cnt = uni(e.cnt);
    for (uint32_t i = first; i < first + cnt; i++) {
      const CsgPrim pr = sload_record(B.prims + i);
      const bool    pn = box_d2<R>(L.x, L.y, L.z, pr.bv[0], pr.bv[1], pr.bv[2], pr.bv[3], pr.bv[4], pr.bv[5]) <= bound;
      if (ballot(pn) == 0) continue;
      const int base = L.vtop;
      for (uint32_t k = uni(pr.start); k < uni(pr.start) + uni(pr.count); k++)
        csg_step<R, KM, NB>(g, fetch<R>(g, k), L, pn);
      const R v = vat(L, base);
      L.vtop    = base;
      if (pn) {
        if (smooth) {
          if (v < a) { b = a; a = v; }
          else if (v < b) b = v;
        }
        else a = min(a, v);
      }
    }
    if (active) {
      const R c = max(R(0), smooth ? b : a);
      bound     = c * c;
    }
  }
  vat(L, L.vtop++) = smooth ? smooth_min<R>(a, b, q.p[0], q.n) : a;
}

template <class R, int KM, int NB>
__device__ __forceinline__ R csg_run(const CsgDev& g, CsgLane<R>& L, bool active)
{
  for (uint32_t i = 0; i < g.main_count; i++) {
    const CsgParams<R> q = fetch<R>(g, i);
    if (q.op == kCsgBvhUnion || q.op == kCsgBvhSmoothUnion) csg_bvh<R, KM, NB>(g, q, L, active);
    else csg_step<R, KM, NB>(g, q, L, active);
  }
  return vat(L, 0);
}

template <int KM, int NB>
__global__ __launch_bounds__(256) void csg32_kernel(const CsgDev g, const QueryArgs a)
{
  __shared__ CsgLds<float, 256> lds;
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  const uint64_t n   = a.n;
  if ((uint64_t)blk * 256u >= n) return; // whole workgroup
  const uint64_t tid    = (uint64_t)blk * 256u + threadIdx.x;
  const bool     active = tid < n;
  uint32_t       qi     = 0;
  CsgLane<float> L{0.0f, 0.0f, 0.0f, 0, 0, &lds.v[0][threadIdx.x], &lds.p[0][0][threadIdx.x], 256u};
  if (active) {
    qi  = a.idx ? a.idx[tid] : (uint32_t)tid;
    L.x = a.pts[3 * (size_t)qi];
    L.y = a.pts[3 * (size_t)qi + 1];
    L.z = a.pts[3 * (size_t)qi + 2];
  }
  const float v = csg_run<float, KM, NB>(g, L, active);
  if (active) a.out_d[qi] = v;
  if (a.recheck_list) wave_append(active && fabsf(v) < a.recheck_abs, qi, a.recheck_list, a.recheck_count);
}

// FP64 pass over the device-side recheck list (length *a.count_ptr): a fixed grid strides over it.
template <int KM, int NB>
__global__ __launch_bounds__(128) void csg64_kernel(const CsgDev g, const QueryArgs a)
{
  __shared__ CsgLds<double, 128> lds;
  const uint64_t n     = *a.count_ptr;
  const uint64_t waves = (uint64_t)gridDim.x * 2u;
  for (uint64_t w = uni(blockIdx.x * 2u + (threadIdx.x >> 6)); w * 64u < n; w += waves) {
    const uint64_t  tid    = w * 64u + __lane_id();
    const bool      active = tid < n;
    uint32_t        qi     = 0;
    CsgLane<double> L{0.0, 0.0, 0.0, 0, 0, &lds.v[0][threadIdx.x], &lds.p[0][0][threadIdx.x], 128u};
    if (active) {
      qi  = a.idx[tid];
      L.x = a.pts[3 * (size_t)qi];
      L.y = a.pts[3 * (size_t)qi + 1];
      L.z = a.pts[3 * (size_t)qi + 2];
    }
    const double v = csg_run<double, KM, NB>(g, L, active);
    if (active) a.out_d[qi] = (float)v;
  }
}

template <int KM, int NB>
hipError_t launch32(const CsgDev& g, const QueryArgs& a, hipStream_t s)
{
  const uint32_t blocks = (uint32_t)((a.n + 255) / 256);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((csg32_kernel<KM, NB>), dim3(blocks), dim3(256), 0, s, g, a);
  return hipGetLastError();
}

template <int KM, int NB>
hipError_t launch64(const CsgDev& g, const QueryArgs& a, hipStream_t s)
{
  if (!a.count_ptr || !a.idx) return hipErrorInvalidValue;
  const uint32_t blocks = std::min<uint32_t>((uint32_t)((a.n + 127) / 128), 2048u);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((csg64_kernel<KM, NB>), dim3(blocks), dim3(128), 0, s, g, a);
  return hipGetLastError();
}

} // namespace

hipError_t launch_csg32(const CsgDev& g, int KM, int banks, const QueryArgs& a, hipStream_t s)
{
  if (KM == 4) {
    if (banks <= 1) return launch32<4, 1>(g, a, s);
    if (banks == 2) return launch32<4, 2>(g, a, s);
    return launch32<4, 4>(g, a, s);
  }
  if (banks <= 1) return launch32<8, 1>(g, a, s);
  if (banks == 2) return launch32<8, 2>(g, a, s);
  return launch32<8, 4>(g, a, s);
}

hipError_t launch_csg64(const CsgDev& g, int KM, int banks, const QueryArgs& a, hipStream_t s)
{
  if (KM == 4) {
    if (banks <= 1) return launch64<4, 1>(g, a, s);
    if (banks == 2) return launch64<4, 2>(g, a, s);
    return launch64<4, 4>(g, a, s);
  }
  if (banks <= 1) return launch64<8, 1>(g, a, s);
  if (banks == 2) return launch64<8, 2>(g, a, s);
  return launch64<8, 4>(g, a, s);
}

} // namespace ebgpu
