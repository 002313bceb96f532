// Device-side building blocks of the packet traversal, shared by the mesh, scene and batched
// kernels (included by their .hip translation units; everything here is inline or a template).
#pragma once

#include "ebgpu_internal.hpp"

#include <hip/hip_runtime.h>

namespace ebgpu {
namespace dev {

// ---------------------------------------------------------------------------------------
// wavefront helpers
// ---------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }
__device__ __forceinline__ float rdlanef(float v, int l)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
// Lane write of a wave-uniform value into lane l (wave-uniform): one v_writelane per field.
// The compiler's writelane intrinsic (this clang has no builtin for it), so that hazards and
// register classes stay the compiler's business.
__device__ int amdgcn_writelane(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ uint32_t wrlane(uint32_t v, int l, uint32_t old)
{
  return (uint32_t)amdgcn_writelane((int)v, l, (int)old);
}
__device__ __forceinline__ float wrlanef(float v, int l, float old)
{
  return __int_as_float(amdgcn_writelane(__float_as_int(v), l, __float_as_int(old)));
}
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }

// Lane mask of a predicate, taken straight from the compare (no round trip through an int).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Number of set bits of m below this lane, plus base.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m, uint32_t base)
{
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, base));
}

// Hand-off of LDS data between the lanes of one wavefront: a wavefront's LDS instructions
// complete in order, so only the compiler has to be kept from moving accesses across this point.
__device__ __forceinline__ void wave_lds_handoff()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Minimum / maximum over the 64 lanes of an unsigned value (all lanes active): DPP within rows
// of 16, then the four row results combined in scalar registers.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v)
{
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));  // quad_perm 1,0,3,2
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));  // quad_perm 2,3,0,1
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false)); // row_half_mirror
  v = min(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false)); // row_mirror
  const uint32_t a = rdlane(v, 0), b = rdlane(v, 16), c = rdlane(v, 32), d = rdlane(v, 48);
  return min(min(a, b), min(c, d));
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v)
{
  v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, false));
  v = max(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, false));
  const uint32_t a = rdlane(v, 0), b = rdlane(v, 16), c = rdlane(v, 32), d = rdlane(v, 48);
  return max(max(a, b), max(c, d));
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v)
{
  // few calls per wave: plain butterfly through ds_swizzle-free shuffles
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Hardware dispatches workgroup b to XCD b % 8. Give each XCD one contiguous range of the
// (Morton-sorted) queries so that its L2 holds one region of the tree.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
  const uint32_t x = b & 7u, i = b >> 3, per = nb >> 3, rem = nb & 7u;
  return x * per + min(x, rem) + i;
}

// ---------------------------------------------------------------------------------------
// wave-uniform record loads
// ---------------------------------------------------------------------------------------

template <int N>
struct SWords
{
  uint32_t w[N];
};

// Load N dwords (16-byte aligned, wave-uniform address) of read-only data into SGPRs: every
// load is issued back to back and waited for once. The loads go through the constant address
// space (the tree and triangle arrays are not written while a query kernel runs), so they are
// scalar loads regardless of stores elsewhere in the kernel; the empty asm statements take each
// value as an SGPR operand at this point, which keeps the compiler from sinking the loads into
// the branches that use them (each sunk load would cost its own memory round trip).
template <int N>
__device__ __forceinline__ SWords<N> sload(const void* p)
{
  static_assert(N % 4 == 0, "sload: whole 16-byte units");
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(4))) u32x4 ConstU4;
  SWords<N>      r;
  ConstU4*       q = (ConstU4*)p;
#pragma unroll
  for (int i = 0; i < N / 4; i++) {
    const u32x4 v = q[i];
    r.w[4 * i]     = v.x;
    r.w[4 * i + 1] = v.y;
    r.w[4 * i + 2] = v.z;
    r.w[4 * i + 3] = v.w;
  }
#pragma unroll
  for (int i = 0; i < N; i += 4)
    asm volatile("" ::"s"(r.w[i]), "s"(r.w[i + 1]), "s"(r.w[i + 2]), "s"(r.w[i + 3]));
  return r;
}

template <class T, int N>
__device__ __forceinline__ T as_record(const SWords<N>& s)
{
  static_assert(sizeof(T) == 4 * N, "as_record: size mismatch");
  T t;
  __builtin_memcpy(&t, s.w, sizeof(T));
  return t;
}

template <class T>
__device__ __forceinline__ T sload_record(const T* p)
{
  return as_record<T>(sload<sizeof(T) / 4>(p));
}

// ---------------------------------------------------------------------------------------
// VGPR-lane stack
// ---------------------------------------------------------------------------------------

// A stack entry's count word: the leaf's triangle count (0 for an interior node) in the low bits,
// the mesh walk's bound stamp above them.
constexpr uint32_t kStampShift = 8;
constexpr uint32_t kCountMask  = (1u << kStampShift) - 1u;
static_assert(kMaxLeaf <= (int)kCountMask, "leaf counts must fit below the stamp");

struct Entry
{
  uint32_t ref, cnt;
  float    lx, ly, lz, hx, hy, hz;
};

template <int NB>
struct LaneStack
{
  uint32_t ref[NB], cnt[NB];
  float    lx[NB], ly[NB], lz[NB], hx[NB], hy[NB], hz[NB];

  __device__ __forceinline__ void init()
  {
#pragma unroll
    for (int i = 0; i < NB; i++) {
      ref[i] = cnt[i] = 0u;
      lx[i] = ly[i] = lz[i] = hx[i] = hy[i] = hz[i] = 0.0f;
    }
  }

  // Write entry e (wave-uniform) at stack position pos (wave-uniform).
  __device__ __forceinline__ void put(int pos, const Entry& e)
  {
    const int l = pos & 63;
#pragma unroll
    for (int i = 0; i < NB; i++) {
      if ((pos >> 6) == i) {
        ref[i] = wrlane(e.ref, l, ref[i]);
        cnt[i] = wrlane(e.cnt, l, cnt[i]);
        lx[i]  = wrlanef(e.lx, l, lx[i]);
        ly[i]  = wrlanef(e.ly, l, ly[i]);
        lz[i]  = wrlanef(e.lz, l, lz[i]);
        hx[i]  = wrlanef(e.hx, l, hx[i]);
        hy[i]  = wrlanef(e.hy, l, hy[i]);
        hz[i]  = wrlanef(e.hz, l, hz[i]);
      }
    }
  }

  // The reference and count words of entry pos alone (wave-uniform).
  __device__ __forceinline__ uint32_t ref_at(int pos) const
  {
    uint32_t v = 0u;
#pragma unroll
    for (int i = 0; i < NB; i++)
      if ((pos >> 6) == i) v = rdlane(ref[i], pos & 63);
    return v;
  }
  __device__ __forceinline__ uint32_t cnt_at(int pos) const
  {
    uint32_t v = 0u;
#pragma unroll
    for (int i = 0; i < NB; i++)
      if ((pos >> 6) == i) v = rdlane(cnt[i], pos & 63);
    return v;
  }

  __device__ __forceinline__ Entry get(int pos) const
  {
    const int l = pos & 63;
    Entry     e{};
#pragma unroll
    for (int i = 0; i < NB; i++) {
      if ((pos >> 6) == i) {
        e.ref = rdlane(ref[i], l);
        e.cnt = rdlane(cnt[i], l);
        e.lx  = rdlanef(lx[i], l);
        e.ly  = rdlanef(ly[i], l);
        e.lz  = rdlanef(lz[i], l);
        e.hx  = rdlanef(hx[i], l);
        e.hy  = rdlanef(hy[i], l);
        e.hz  = rdlanef(hz[i], l);
      }
    }
    return e;
  }
};

// ---------------------------------------------------------------------------------------
// distance kernels
// ---------------------------------------------------------------------------------------

template <class R>
__device__ __forceinline__ R box_d2(R px, R py, R pz, float lx, float ly, float lz, float hx, float hy, float hz)
{
  if constexpr (sizeof(R) == 4) {
    // p minus its clamp to the box (v_med3_f32): zero inside, the gap outside
    const float dx = px - __builtin_amdgcn_fmed3f(px, lx, hx);
    const float dy = py - __builtin_amdgcn_fmed3f(py, ly, hy);
    const float dz = pz - __builtin_amdgcn_fmed3f(pz, lz, hz);
    return dx * dx + dy * dy + dz * dz;
  }
  else {
    const R dx = max(R(0), max(R(lx) - px, px - R(hx)));
    const R dy = max(R(0), max(R(ly) - py, py - R(hy)));
    const R dz = max(R(0), max(R(lz) - pz, pz - R(hz)));
    return dx * dx + dy * dy + dz * dz;
  }
}

// Squared distance from p to one triangle (FP32, precomputed record) and, with FEAT, its
// closest feature. The traversal evaluates candidates with FEAT = false (the distance is all
// it compares) and classifies the winner once at the end with FEAT = true; both compute d2 with
// the same operations, so the classification belongs to the distance that won.
template <bool FEAT>
__device__ __forceinline__ void tri_eval32(const TriRec& r, float px, float py, float pz, float& d2, uint32_t& feat)
{
  const float p1x = px - r.v0[0], p1y = py - r.v0[1], p1z = pz - r.v0[2];
  const float p2x = p1x - r.e0[0], p2y = p1y - r.e0[1], p2z = p1z - r.e0[2];
  const float p3x = p1x + r.e2[0], p3y = p1y + r.e2[1], p3z = p1z + r.e2[2];

  const bool  s0 = (r.c0[0] * p1x + r.c0[1] * p1y + r.c0[2] * p1z) > 0.0f;
  const bool  s1 = (r.c1[0] * p2x + r.c1[1] * p2y + r.c1[2] * p2z) > 0.0f;
  const bool  s2 = (r.c2[0] * p3x + r.c2[1] * p3y + r.c2[2] * p3z) > 0.0f;
  const bool  inface = (s0 == s1) & (s1 == s2); // no branch: all three compares are needed anyway
  const float fd     = r.n[0] * p1x + r.n[1] * p1y + r.n[2] * p1z;

  // Outside the face region: the closest point of each edge at its clamped parameter (t <= 0
  // gives the edge's first vertex, t >= 1 its second). The minimum over the three edges is the
  // reference's minimum over the three vertices and the edge interiors with t in (0,1); at the
  // clamps the same vertex vectors p1, p2, p3 are formed, so the distances agree bit for bit
  // (up to the rounding of p3 - e2 and p2 - e1 against p1 and p3).
  const float t0  = __builtin_amdgcn_fmed3f((p1x * r.e0[0] + p1y * r.e0[1] + p1z * r.e0[2]) * r.inv[0], 0.0f, 1.0f);
  const float t1  = __builtin_amdgcn_fmed3f((p2x * r.e1[0] + p2y * r.e1[1] + p2z * r.e1[2]) * r.inv[1], 0.0f, 1.0f);
  const float t2  = __builtin_amdgcn_fmed3f((p3x * r.e2[0] + p3y * r.e2[1] + p3z * r.e2[2]) * r.inv[2], 0.0f, 1.0f);
  const float y0x = p1x - t0 * r.e0[0], y0y = p1y - t0 * r.e0[1], y0z = p1z - t0 * r.e0[2];
  const float y1x = p2x - t1 * r.e1[0], y1y = p2y - t1 * r.e1[1], y1z = p2z - t1 * r.e1[2];
  const float y2x = p3x - t2 * r.e2[0], y2y = p3y - t2 * r.e2[1], y2z = p3z - t2 * r.e2[2];
  const float q0  = y0x * y0x + y0y * y0y + y0z * y0z;
  const float q1  = y1x * y1x + y1y * y1y + y1z * y1z;
  const float q2  = y2x * y2x + y2y * y2y + y2z * y2z;
  // minimum of the three (a NaN from a zero-length edge loses to a number)
  const float r2  = fminf(fminf(q0, q1), q2);
  d2              = inface ? fd * fd : r2;
  if constexpr (FEAT) {
    // the first edge attaining the minimum, and its clamped parameter
    const uint32_t e  = q0 == r2 ? 0u : (q1 == r2 ? 1u : 2u);
    const float    te = e == 0 ? t0 : (e == 1 ? t1 : t2);
    const uint32_t f  = te <= 0.0f ? kVert0 + e : (te >= 1.0f ? kVert0 + (e == 2 ? 0u : e + 1u) : kEdge0 + e);
    feat              = inface ? (uint32_t)kFace : f;
  }
}

// Squared distance from p to a triangle by its in-plane frame (FrameRec): p - A projected on
// (ex, ey, ez) gives (u, v, w); d^2 = w^2 plus, outside the face, the smallest squared distance
// in the plane from (u, v) to the edges A -> B (u clamped to [0, a]), B -> C and A -> C (their
// parameters clamped to [0, 1]). Inside the face means strictly left of all three edges of
// A = (0, 0), B = (a, 0), C = (b, c), c >= 0; on an edge line both forms give the same distance.
// The same quantity as tri_eval32's d2 up to rounding, in 40 instead of 63 VALU operations
// (error relative to |p - A| + |d| below 2^-22 on random, sliver, collinear and point triangles).
__device__ __forceinline__ float frame_eval32(const FrameRec& r, float px, float py, float pz)
{
  const float p1x = px - r.o[0], p1y = py - r.o[1], p1z = pz - r.o[2];
  const float u   = r.ex[0] * p1x + r.ex[1] * p1y + r.ex[2] * p1z;
  const float v   = r.ey[0] * p1x + r.ey[1] * p1y + r.ey[2] * p1z;
  const float w   = r.ez[0] * p1x + r.ez[1] * p1y + r.ez[2] * p1z;
  // inside: the smallest of the three edge-line values is positive (one v_min3 and one compare)
  const bool  inface = fminf(fminf(v, fmaf(r.c, u, -(r.b * v))), fmaf(r.bma, v, fmaf(-r.c, u, r.ca))) > 0.0f;
  const float du  = u - __builtin_amdgcn_fmed3f(u, 0.0f, r.a);
  const float q0  = du * du + v * v;
  const float t1  = __builtin_amdgcn_fmed3f(fmaf(r.k1u, u, fmaf(r.k1v, v, r.k1c)), 0.0f, 1.0f);
  const float y1x = u - fmaf(t1, r.bma, r.a), y1y = fmaf(-t1, r.c, v);
  const float q1  = y1x * y1x + y1y * y1y;
  const float t2  = __builtin_amdgcn_fmed3f(fmaf(r.k2u, u, r.k2v * v), 0.0f, 1.0f);
  const float y2x = fmaf(-t2, r.b, u), y2y = fmaf(-t2, r.c, v);
  const float q2  = y2x * y2x + y2y * y2y;
  return fmaf(w, w, inface ? 0.0f : fminf(fminf(q0, q1), q2));
}

// tri_eval32<false> behind a wavefront-level cull: d^2 = fd^2 + (in-plane distance)^2, and the
// in-plane distance is at least the distance past any edge line the projected point lies beyond
// (c_i points out of the triangle, |c_i| = |e_i|), so lb = fd^2 + max_i max(c_i.p_i, 0)^2 / |e_i|^2
// <= d^2. When no lane's lb reaches its limit `cull` the wavefront skips the rest of the
// evaluation (returns false); otherwise d2 is tri_eval32's, operation for operation. `cull` is
// cull_band(bound): lb and the evaluated d2 round differently (each within 2^-21 (d + 2 hmax) of
// the exact distance), and with the band no triangle whose evaluated d2 would reach the bound is
// skipped, so the result does not depend on the order in which the bound tightened.
__device__ __forceinline__ float cull_band(float bound, float hmax)
{
  const float r = fmaf(__builtin_amdgcn_sqrtf(bound), 1.0f + 0x1p-19f, hmax * 0x1p-18f);
  return r * r * (1.0f + 0x1p-20f);
}

__device__ __forceinline__ bool tri_eval32_culled(const TriRec& r, float px, float py, float pz, float cull, float& d2)
{
  const float p1x = px - r.v0[0], p1y = py - r.v0[1], p1z = pz - r.v0[2];
  const float p2x = p1x - r.e0[0], p2y = p1y - r.e0[1], p2z = p1z - r.e0[2];
  const float p3x = p1x + r.e2[0], p3y = p1y + r.e2[1], p3z = p1z + r.e2[2];
  const float s0  = r.c0[0] * p1x + r.c0[1] * p1y + r.c0[2] * p1z;
  const float s1  = r.c1[0] * p2x + r.c1[1] * p2y + r.c1[2] * p2z;
  const float s2  = r.c2[0] * p3x + r.c2[1] * p3y + r.c2[2] * p3z;
  const float fd  = r.n[0] * p1x + r.n[1] * p1y + r.n[2] * p1z;
  const float o0 = fmaxf(s0, 0.0f), o1 = fmaxf(s1, 0.0f), o2 = fmaxf(s2, 0.0f);
  const float out = fmaxf(fmaxf(o0 * o0 * r.inv[0], o1 * o1 * r.inv[1]), o2 * o2 * r.inv[2]);
  const float lb  = fmaf(fd, fd, out);
  if (ballot(lb <= cull) == 0) return false;
  const bool  inface = ((s0 > 0.0f) == (s1 > 0.0f)) && ((s1 > 0.0f) == (s2 > 0.0f));
  const float t0  = __builtin_amdgcn_fmed3f((p1x * r.e0[0] + p1y * r.e0[1] + p1z * r.e0[2]) * r.inv[0], 0.0f, 1.0f);
  const float t1  = __builtin_amdgcn_fmed3f((p2x * r.e1[0] + p2y * r.e1[1] + p2z * r.e1[2]) * r.inv[1], 0.0f, 1.0f);
  const float t2  = __builtin_amdgcn_fmed3f((p3x * r.e2[0] + p3y * r.e2[1] + p3z * r.e2[2]) * r.inv[2], 0.0f, 1.0f);
  const float y0x = p1x - t0 * r.e0[0], y0y = p1y - t0 * r.e0[1], y0z = p1z - t0 * r.e0[2];
  const float y1x = p2x - t1 * r.e1[0], y1y = p2y - t1 * r.e1[1], y1z = p2z - t1 * r.e1[2];
  const float y2x = p3x - t2 * r.e2[0], y2y = p3y - t2 * r.e2[1], y2z = p3z - t2 * r.e2[2];
  const float q0  = y0x * y0x + y0y * y0y + y0z * y0z;
  const float q1  = y1x * y1x + y1y * y1y + y1z * y1z;
  const float q2  = y2x * y2x + y2y * y2y + y2z * y2z;
  d2              = inface ? fd * fd : fminf(fminf(q0, q1), q2);
  return true;
}

// Signed distance of the chosen feature (FP32): face -> n.(p-v0); vertex/edge ->
// sqrt(d2) * sgn(pseudo-normal . (p - closest point)), sgn(x) = x > 0 ? 1 : -1.
__device__ __forceinline__ float resolve32(const MeshDev& m, uint32_t t, uint32_t feat, float d2, float px, float py, float pz)
{
  const TriRec& r   = m.tris[t];
  const float   p1x = px - r.v0[0], p1y = py - r.v0[1], p1z = pz - r.v0[2];
  if (feat == kFace) return r.n[0] * p1x + r.n[1] * p1y + r.n[2] * p1z;
  float vx, vy, vz;
  const float* nrm;
  if (feat <= kVert2) {
    const uint32_t vid = m.feat[8 * t + (feat - kVert0)];
    nrm                = m.vn32 + 4 * (size_t)vid;
    if (feat == kVert0) { vx = p1x; vy = p1y; vz = p1z; }
    else if (feat == kVert1) { vx = p1x - r.e0[0]; vy = p1y - r.e0[1]; vz = p1z - r.e0[2]; }
    else { vx = p1x + r.e2[0]; vy = p1y + r.e2[1]; vz = p1z + r.e2[2]; }
  }
  else {
    const uint32_t e  = feat - kEdge0;
    const uint32_t he = m.feat[8 * t + 3 + e];
    nrm               = m.en32 + 4 * (size_t)he;
    float ax, ay, az, ex, ey, ez, inv;
    if (e == 0) { ax = p1x; ay = p1y; az = p1z; ex = r.e0[0]; ey = r.e0[1]; ez = r.e0[2]; inv = r.inv[0]; }
    else if (e == 1) {
      ax = p1x - r.e0[0]; ay = p1y - r.e0[1]; az = p1z - r.e0[2];
      ex = r.e1[0]; ey = r.e1[1]; ez = r.e1[2]; inv = r.inv[1];
    }
    else {
      ax = p1x + r.e2[0]; ay = p1y + r.e2[1]; az = p1z + r.e2[2];
      ex = r.e2[0]; ey = r.e2[1]; ez = r.e2[2]; inv = r.inv[2];
    }
    const float tt = (ax * ex + ay * ey + az * ez) * inv;
    vx             = ax - tt * ex;
    vy             = ay - tt * ey;
    vz             = az - tt * ez;
  }
  const float s = (nrm[0] * vx + nrm[1] * vy + nrm[2] * vz) > 0.0f ? 1.0f : -1.0f;
  return s * sqrtf(d2);
}

// FP64: one triangle with the reference lane formulas on the exact (float) vertex positions
// and the FP64 pseudo-normals. Returns d^2 (face: (n.p1)^2) and the feature.
__device__ __forceinline__ void tri_eval64(const float* pos, const double* fn, double px, double py, double pz,
                                           double& d2, uint32_t& feat)
{
  const double v0x = pos[0], v0y = pos[1], v0z = pos[2];
  const double v1x = pos[3], v1y = pos[4], v1z = pos[5];
  const double v2x = pos[6], v2y = pos[7], v2z = pos[8];
  const double nx = fn[0], ny = fn[1], nz = fn[2];
  const double ax = v1x - v0x, ay = v1y - v0y, az = v1z - v0z; // v21
  const double bx = v2x - v1x, by = v2y - v1y, bz = v2z - v1z; // v32
  const double cx = v0x - v2x, cy = v0y - v2y, cz = v0z - v2z; // v13
  const double p1x = px - v0x, p1y = py - v0y, p1z = pz - v0z;
  const double p2x = px - v1x, p2y = py - v1y, p2z = pz - v1z;
  const double p3x = px - v2x, p3y = py - v2y, p3z = pz - v2z;
  // (v x n) . p
  const double s0 = (ay * nz - az * ny) * p1x + (az * nx - ax * nz) * p1y + (ax * ny - ay * nx) * p1z;
  const double s1 = (by * nz - bz * ny) * p2x + (bz * nx - bx * nz) * p2y + (bx * ny - by * nx) * p2z;
  const double s2 = (cy * nz - cz * ny) * p3x + (cz * nx - cx * nz) * p3y + (cx * ny - cy * nx) * p3z;
  const int    ss = (s0 > 0.0 ? 1 : -1) + (s1 > 0.0 ? 1 : -1) + (s2 > 0.0 ? 1 : -1);
  if (ss >= 2 || ss <= -2) {
    const double fd = nx * p1x + ny * p1y + nz * p1z;
    d2              = fd * fd;
    feat            = kFace;
    return;
  }
  double   r2 = p1x * p1x + p1y * p1y + p1z * p1z;
  uint32_t f  = kVert0;
  double   q  = p2x * p2x + p2y * p2y + p2z * p2z;
  if (q < r2) { r2 = q; f = kVert1; }
  q = p3x * p3x + p3y * p3y + p3z * p3z;
  if (q < r2) { r2 = q; f = kVert2; }
  double t = (p1x * ax + p1y * ay + p1z * az) / (ax * ax + ay * ay + az * az);
  if (t > 0.0 && t < 1.0) {
    const double yx = p1x - t * ax, yy = p1y - t * ay, yz = p1z - t * az;
    q               = yx * yx + yy * yy + yz * yz;
    if (q < r2) { r2 = q; f = kEdge0; }
  }
  t = (p2x * bx + p2y * by + p2z * bz) / (bx * bx + by * by + bz * bz);
  if (t > 0.0 && t < 1.0) {
    const double yx = p2x - t * bx, yy = p2y - t * by, yz = p2z - t * bz;
    q               = yx * yx + yy * yy + yz * yz;
    if (q < r2) { r2 = q; f = kEdge1; }
  }
  t = (p3x * cx + p3y * cy + p3z * cz) / (cx * cx + cy * cy + cz * cz);
  if (t > 0.0 && t < 1.0) {
    const double yx = p3x - t * cx, yy = p3y - t * cy, yz = p3z - t * cz;
    q               = yx * yx + yy * yy + yz * yz;
    if (q < r2) { r2 = q; f = kEdge2; }
  }
  d2   = r2;
  feat = f;
}

__device__ __forceinline__ double resolve64(const MeshDev& m, uint32_t t, uint32_t feat, double d2, double px,
                                            double py, double pz)
{
  const float*  pos = m.tripos + 12 * (size_t)t;
  const double* fn  = m.fn64 + 4 * (size_t)t;
  const double  v0x = pos[0], v0y = pos[1], v0z = pos[2];
  const double  p1x = px - v0x, p1y = py - v0y, p1z = pz - v0z;
  if (feat == kFace) return fn[0] * p1x + fn[1] * p1y + fn[2] * p1z;
  double        vx, vy, vz;
  const double* nrm;
  if (feat <= kVert2) {
    const uint32_t k = feat - kVert0;
    nrm              = m.vn64 + 4 * (size_t)m.feat[8 * t + k];
    vx               = px - double(pos[3 * k]);
    vy               = py - double(pos[3 * k + 1]);
    vz               = pz - double(pos[3 * k + 2]);
  }
  else {
    const uint32_t e = feat - kEdge0, a = e, b = (e + 1) % 3;
    nrm              = m.en64 + 4 * (size_t)m.feat[8 * t + 3 + e];
    const double ax = double(pos[3 * b]) - double(pos[3 * a]), ay = double(pos[3 * b + 1]) - double(pos[3 * a + 1]),
                 az = double(pos[3 * b + 2]) - double(pos[3 * a + 2]);
    const double qx = px - double(pos[3 * a]), qy = py - double(pos[3 * a + 1]), qz = pz - double(pos[3 * a + 2]);
    const double tt = (qx * ax + qy * ay + qz * az) / (ax * ax + ay * ay + az * az);
    vx              = qx - tt * ax;
    vy              = qy - tt * ay;
    vz              = qz - tt * az;
  }
  const double s = (nrm[0] * vx + nrm[1] * vy + nrm[2] * vz) > 0.0 ? 1.0 : -1.0;
  return sqrt(d2) * s;
}

// Wave-aggregated append of `idx` for the lanes where pred holds.
__device__ __forceinline__ void wave_append(bool pred, uint32_t idx, uint32_t* list, uint32_t* count)
{
  const uint64_t mask = ballot(pred);
  if (mask == 0) return;
  const int      lane   = __lane_id();
  const uint32_t n      = (uint32_t)__popcll(mask);
  const int      leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t       base   = 0;
  if (lane == leader) base = atomicAdd(count, n);
  base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
  if (pred) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = idx;
}

// ---------------------------------------------------------------------------------------
// closest-face candidates (exact closest face from an FP32 walk)
// ---------------------------------------------------------------------------------------

// The FP32 distance of a point to a triangle differs from the FP64 lane formula's by at most
// E(d) = 2^-21 (d + 2 hmax) (d the distance, hmax the mesh's longest edge: every rounding acts on
// a vector no longer than d + 2 hmax; measured maximum 3 * 2^-24 (d + 2 hmax)). So the face that
// is closest in FP64 has an FP32 distance within 2 E of the FP32 minimum, and a lane that keeps
// every triangle inside that band can decide its face exactly, in FP64, among them.
constexpr int kFaceCands = 8;

struct FaceCands
{
  float    d[kFaceCands]; // FP32 d^2 of the kept triangles, ascending (+inf: empty slot)
  uint32_t t[kFaceCands];
  float    lim;           // band of the current minimum: triangles with d^2 <= lim are kept
  float    evict;         // smallest d^2 pushed out of a full list
  float    habs;          // 2^-19 hmax: the band's absolute part

  __device__ __forceinline__ void init(float hmax)
  {
#pragma unroll
    for (int i = 0; i < kFaceCands; i++) {
      d[i] = INFINITY;
      t[i] = kInvalid;
    }
    lim   = INFINITY;
    evict = INFINITY;
    habs  = hmax * 0x1p-19f;
  }

  // band(d^2) = (d (1 + 2^-20) + 2^-19 hmax)^2, rounded up
  __device__ __forceinline__ float band(float d2) const
  {
    const float r = fmaf(__builtin_amdgcn_sqrtf(d2), 1.0f + 0x1p-20f, habs);
    return r * r * (1.0f + 0x1p-21f);
  }

  // Sorted insert (fixed positions only: the list stays in VGPRs); a value pushed out of the
  // full list is remembered in `evict`.
  __device__ __forceinline__ void insert(float x, uint32_t xt)
  {
#pragma unroll
    for (int i = 0; i < kFaceCands; i++) {
      const bool     sw = x < d[i];
      const float    od = d[i];
      const uint32_t ot = t[i];
      d[i]              = sw ? x : od;
      t[i]              = sw ? xt : ot;
      x                 = sw ? od : x;
      xt                = sw ? ot : xt;
    }
    evict = fminf(evict, x);
  }
};

// ---------------------------------------------------------------------------------------
// the packet traversal
// ---------------------------------------------------------------------------------------

// Interior node: test the K child boxes and push the ones some lane needs, ranked so that the
// farthest lies deepest and the nearest is popped next. The order key is the box distance seen
// by the representative lane `rep` (see rep_lane) for dense, curve-sorted mesh packets: one
// readlane instead of a wavefront minimum (DPP and 4 readlanes per child), and as good an
// order there (16M queries, 327,680-face displaced sphere: 1.6% fewer triangle evaluations, 5.7% less time). With WMIN the key is the
// minimum over the lanes that need the child: better for packets whose lanes are spread (the
// FP64 passes over sparse lists, CSG scenes with primitives the size of a packet). `tag` is OR-ed
// into each pushed count word (the mesh walk's stamp, above the leaf count's bits).
template <int K, int NB, class R, bool WMIN>
__device__ __forceinline__ void push_children(const Node<K>& nd, R px, R py, R pz, R bound, LaneStack<NB>& st,
                                              int& top, int rep, uint32_t tag = 0u)
{
  uint32_t key[K];
  uint64_t msk[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    msk[k] = 0;
    key[k] = 0xFFFFFFFFu;
    if (nd.cnt[k] != kInvalid) {
      const R    d2 = box_d2<R>(px, py, pz, nd.lox[k], nd.loy[k], nd.loz[k], nd.hix[k], nd.hiy[k], nd.hiz[k]);
      const bool nk = d2 <= bound;
      msk[k]        = ballot(nk);
      if (msk[k]) {
        if constexpr (WMIN) key[k] = wave_min_u32(nk ? __float_as_uint(float(d2)) : 0xFFFFFFFFu);
        else key[k] = rdlane(__float_as_uint(float(d2)), rep);
      }
    }
  }
  int pushed = 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    if (!msk[k]) continue;
    int rank = 0;
#pragma unroll
    for (int j = 0; j < K; j++) {
      if (j == k || !msk[j]) continue;
      rank += (key[j] > key[k] || (key[j] == key[k] && j < k)) ? 1 : 0;
    }
    Entry c{nd.ref[k], nd.cnt[k] | tag, nd.lox[k], nd.loy[k], nd.loz[k], nd.hix[k], nd.hiy[k], nd.hiz[k]};
    st.put(top + rank, c);
    pushed++;
  }
  top += pushed;
}

// Representative lane of a packet for the child order: lane 32, the middle of a curve-sorted
// packet, when it is active (bound >= 0), else the first active lane (list-driven passes, where
// the active lanes are sparse).
template <class R>
__device__ __forceinline__ int rep_lane(R bound)
{
  const uint64_t act = ballot(bound >= R(0));
  return ((act >> 32) & 1ull) ? 32 : (act ? __builtin_ctzll(act) : 0);
}

// Per-wavefront LDS of the pair-compacted leaf evaluation (leaf_pairs): the leaf's triangle
// records, every (lane, triangle) pair the screen passed, and each lane's best key.
struct alignas(64) PairLds
{
  float4   rec[kMaxLeaf * 8];   // kMaxLeaf FrameRecs of 128 B
  uint64_t best[64];            // (d^2 bits << 32) | triangle, per lane
  uint16_t pair[kMaxLeaf * 64]; // (lane << 4) | slot in the leaf
};
static_assert(kMaxLeaf <= 16, "leaf slots are 4 bits");
static_assert(sizeof(PairLds) == 4608, "PairLds layout");

// Counting walks (wave-uniform): the pairs the pair-compacted leaves evaluated, their 64-lane
// batches, the culling-record loads of the screen (4 triangles, 128 B each) and the leaves whose
// slab was tested (one 4-B load each).
struct PairCounts
{
  uint32_t pairs, batches, groups, leaves;
};

// FP32 distance-only leaf, compacted: the screen (bounding sphere and plane against each lane's
// bound, as in the wavefront-wide leaf) is run for every triangle of the leaf, and each passing
// (lane, triangle) pair is appended to a list in LDS; then each lane evaluates one pair at a time
// -- its triangle's record read from LDS, its query point from the pair's lane (ds_bpermute) --
// and folds (d^2, triangle) into the pair lane's best key with an LDS atomic minimum. A
// wavefront-wide evaluation costs a full evaluation's instructions whenever one lane needs the
// triangle; here 64 needed pairs share those instructions, whichever lanes and triangles they
// come from. A pair is evaluated on the triangle's in-plane frame (frame_eval32, a third fewer
// instructions than tri_eval32), over the pairs the wavefront-wide leaf would evaluate (the
// screen is the same conservative test, against bounds tightened once per leaf instead of per
// triangle). The walk's d^2 is this function's minimum (among equal d^2 the lower triangle
// index wins), whichever packets and subtree walks the query was part of; it agrees with the
// wavefront-wide leaf's tri_eval32 minimum to FP32 rounding, and where two triangles' distances
// agree to rounding the two may pick different ones.
// Before the screen, the leaf's slab (MeshDev::lthick, about its first triangle's plane) skips
// leaves that no lane comes within. The leaf's records are fetched (one 32-byte piece per lane)
// only when some pair needs them: loading every visited leaf ahead of the screen, to hide the
// load's latency behind it, fetched three leaves for each one used (1.4x the DRAM traffic) and
// was 1% slower.
template <bool COUNT>
__device__ __forceinline__ void leaf_pairs(const MeshDev& m, uint32_t first, uint32_t cnt, float px, float py,
                                           float pz, float& bound, float& best, uint32_t& btri, uint32_t& n_tris,
                                           uint32_t& w_tris, PairLds& L, PairCounts* pc)
{
  const uint32_t lane = __lane_id();
  const uint32_t q0   = 2u * lane; // this lane's 16-byte pieces q0, q0 + 1 of the leaf's records
  const bool     mine = q0 < 8u * cnt;
  const float4* __restrict__ src = reinterpret_cast<const float4*>(m.frames + first);

  const float rb = __builtin_amdgcn_sqrtf(bound) * 1.000001f; // 1-ulp hardware root, rounded up
  SWords<32>  sp = sload<32>(m.tsph + 8 * (size_t)first); // culling records of the first 4 triangles
  if (m.lthick) {
    // the leaf's slab (its first triangle's plane, widened to hold the whole leaf) is a lower
    // bound on every lane's distance to it: the screen is skipped when no lane reaches it. The
    // plane comes with the first culling records (the same round trip).
    typedef const __attribute__((address_space(4))) float ConstF;
    const float t  = *(ConstF*)(m.lthick + first);
    const float pd = fmaf(__uint_as_float(sp.w[4]), px,
                          fmaf(__uint_as_float(sp.w[5]), py, fmaf(__uint_as_float(sp.w[6]), pz, __uint_as_float(sp.w[7]))));
    if (COUNT && pc) {
      pc->leaves++;
      pc->groups++;
    }
    if (ballot(fabsf(pd) <= rb + t) == 0) return;
    if (COUNT && pc) pc->groups += (cnt + 3u) / 4u - 1u;
  }
  else if (COUNT && pc) pc->groups += (cnt + 3u) / 4u;
  uint32_t total = 0; // pairs listed (wave-uniform)
  for (uint32_t c0 = 0; c0 < cnt; c0 += 4) {
    if (c0) sp = sload<32>(m.tsph + 8 * (size_t)(first + c0));
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (c0 + i < cnt) {
        const float sx = px - __uint_as_float(sp.w[8 * i]), sy = py - __uint_as_float(sp.w[8 * i + 1]),
                    sz = pz - __uint_as_float(sp.w[8 * i + 2]);
        const float lim  = rb + __uint_as_float(sp.w[8 * i + 3]);
        const bool  in_s = sx * sx + sy * sy + sz * sz <= lim * lim;
        const float pd   = fmaf(__uint_as_float(sp.w[8 * i + 4]), px,
                                fmaf(__uint_as_float(sp.w[8 * i + 5]), py,
                                     fmaf(__uint_as_float(sp.w[8 * i + 6]), pz, __uint_as_float(sp.w[8 * i + 7]))));
        const bool  in_p = fabsf(pd) <= rb;
        if (COUNT) n_tris += (in_s && in_p) ? 1u : 0u;
        const uint64_t pm = ballot(in_s) & ballot(in_p);
        if (pm) {
          // this lane's bit of pm is in_s && in_p: the write is predicated on the compares
          if (in_s && in_p) L.pair[mbcnt(pm, total)] = (uint16_t)((lane << 4) | (c0 + (uint32_t)i));
          total += (uint32_t)__popcll(pm);
        }
      }
    }
  }
  if (total == 0) return;
  if (COUNT) {
    w_tris += cnt;
    if (pc) {
      pc->pairs += total;
      pc->batches += (total + 63u) / 64u;
    }
  }
  if (mine) {
    L.rec[q0]     = src[q0];
    L.rec[q0 + 1] = src[q0 + 1];
  }
  L.best[lane] = ((uint64_t)__float_as_uint(best) << 32) | btri;
  wave_lds_handoff();
  for (uint32_t b0 = 0; b0 < total; b0 += 64) {
    const uint32_t k  = b0 + lane;
    const bool     ok = k < total;
    uint32_t       e  = lane << 4; // idle lanes: their own point
    if (ok) e = L.pair[k];
    // all 64 lanes take part in the permutes (an inactive source lane would not be read)
    const int   sa = (int)((e >> 4) << 2);
    const float qx = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(px)));
    const float qy = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(py)));
    const float qz = __int_as_float(__builtin_amdgcn_ds_bpermute(sa, __float_as_int(pz)));
    if (ok) {
      const uint32_t slot = e & 15u;
      const float    d2   = frame_eval32(*reinterpret_cast<const FrameRec*>(&L.rec[8 * slot]), qx, qy, qz);
      atomicMin(reinterpret_cast<unsigned long long*>(&L.best[e >> 4]),
                ((unsigned long long)__float_as_uint(d2) << 32) | (unsigned long long)(first + slot));
    }
  }
  wave_lds_handoff();
  const uint64_t kb = L.best[lane];
  best              = __uint_as_float((uint32_t)(kb >> 32));
  btri              = (uint32_t)kb;
  bound             = best < bound ? best : bound; // bound stays min(bound, best)
}

// R = float: FP32 pass on TriRec records. R = double: FP64 recheck on raw positions.
// The walk starts at `start` (the root, or a subtree root of a split packet).
// FACE (FP32 only): the cull bound is the band of the minimum (FaceCands::band) rather than the
// minimum itself, the plane culls are off (their FP32 error is one of the coordinates, not of
// the distance), and every triangle inside the band is kept in fc.
// PAIRS (FP32 distance-only walks): leaves are evaluated by leaf_pairs in the wavefront's PairLds.
template <int K, int NB, class R, bool COUNT, bool WMIN = (sizeof(R) == 8), bool FACE = false, bool PAIRS = false>
__device__ __forceinline__ void traverse_from(const MeshDev& m, const Entry& start, R px, R py, R pz, R bound, R& best,
                                              uint32_t& btri, uint32_t& bfeat, uint32_t& n_inner, uint32_t& n_tris,
                                              uint32_t& w_inner, uint32_t& w_tris, FaceCands* fc = nullptr,
                                              PairLds* pl = nullptr, PairCounts* pc = nullptr)
{
  static_assert(!FACE || sizeof(R) == 4, "closest-face candidates: FP32 walk only");
  static_assert(!PAIRS || (sizeof(R) == 4 && !FACE), "pair-compacted leaves: FP32 distance-only walks");
  const Node<K>* __restrict__ nodes = static_cast<const Node<K>*>(m.nodes);
  LaneStack<NB>               st;
  st.init();
  const int lane = __lane_id();
  const int rep  = rep_lane(bound);
  int       top  = 0;
  st.put(top++, start);
  // Stamp of the bounds: it advances after every leaf that tightened some lane's bound, and each
  // pushed entry carries it above the count's bits. An entry whose stamp is current was pushed
  // because some lane's box distance^2 was within its bound, and no bound has changed since, so
  // the test at the pop would pass again: it is skipped (the box fields are not even read). A
  // skipped test can only cost work, never change a result; the counting walk always tests.
  uint32_t stamp = 1u; // the start entry carries 0: always tested

  while (top > 0) {
    --top;
    const uint32_t ecnt = st.cnt_at(top);
    bool           need = true;
    if (COUNT || (ecnt >> kStampShift) != stamp) {
      const Entry e   = st.get(top);
      const R     d2e = box_d2<R>(px, py, pz, e.lx, e.ly, e.lz, e.hx, e.hy, e.hz);
      need            = d2e <= bound;
      if (ballot(need) == 0) continue;
    }
    const uint32_t eref = st.ref_at(top), cnt = ecnt & kCountMask;

    if (cnt == 0u) {
      if (COUNT) { n_inner += need ? 1u : 0u; w_inner++; }
      push_children<K, NB, R, WMIN>(sload_record(nodes + uni(eref)), px, py, pz, bound, st, top, rep,
                                    stamp << kStampShift);
      // The capacity is the tree's exact worst case, computed at build time; should a walk ever
      // exceed it, the pushes beyond it were dropped: record that in the mesh's status word
      // (the host reports it) and go on with the entries that are held.
      if (top > 64 * NB) {
        if (lane == 0 && m.status) atomicOr(m.status, kStatusStackOverflow);
        top = 64 * NB;
      }
    }
    else if constexpr (PAIRS) {
      const R bound0 = bound;
      leaf_pairs<COUNT>(m, uni(eref), cnt, px, py, pz, bound, best, btri, n_tris, w_tris, *pl, pc);
      if (ballot(bound < bound0)) stamp++;
    }
    else {
      // leaf: a triangle is evaluated by the whole wavefront (keeping the minimum is exact for
      // every lane) unless, for every lane, its bounding sphere or its plane lies beyond the
      // lane's bound (both are lower bounds on the distance to the triangle). Inactive lanes
      // have bound < 0, so rb is NaN and they never ask for a triangle. The culling data of 4
      // triangles arrives in one scalar round trip (the array is padded past the last leaf),
      // then each needed triangle record in one more, where a lower bound (the plane and the
      // edge lines, tri_eval32_culled) is tested against the bound as tightened by the leaf's
      // earlier triangles.
      const uint32_t first  = uni(eref);
      const R        bound0 = bound;
      R              rb;
      [[maybe_unused]] float cull = 0.0f; // FP32 distance-only walks: the lower-bound cull's limit
      if constexpr (sizeof(R) == 4) {
        rb = __builtin_amdgcn_sqrtf(bound) * 1.000001f; // 1-ulp hardware root, rounded up
        if constexpr (!FACE) cull = cull_band(bound, m.hmax);
      }
      else rb = sqrt(bound);
      for (uint32_t c0 = 0; c0 < cnt; c0 += 4) {
        const SWords<32> sp   = sload<32>(m.tsph + 8 * (size_t)(first + c0));
        uint32_t         need = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          if (c0 + i < cnt) {
            const R    sx  = px - R(__uint_as_float(sp.w[8 * i])), sy = py - R(__uint_as_float(sp.w[8 * i + 1])),
                       sz  = pz - R(__uint_as_float(sp.w[8 * i + 2]));
            const R    lim = rb + R(__uint_as_float(sp.w[8 * i + 3]));
            const bool in_s = sx * sx + sy * sy + sz * sz <= lim * lim;
            bool       in_p = true;
            if constexpr (sizeof(R) == 4 && !FACE) {
              // the FP32 plane (rounded normal and offset) may misjudge by a few ulps of the
              // coordinates; the FP64 pass, whose seed only just contains the minimum, skips it
              const float pd = fmaf(__uint_as_float(sp.w[8 * i + 4]), px,
                                    fmaf(__uint_as_float(sp.w[8 * i + 5]), py,
                                         fmaf(__uint_as_float(sp.w[8 * i + 6]), pz, __uint_as_float(sp.w[8 * i + 7]))));
              in_p = fabsf(pd) <= rb;
            }
            if (COUNT) n_tris += (in_s && in_p) ? 1u : 0u;
            // two ballots straight from the compares (a ballot of their conjunction went
            // through a VGPR)
            need |= (ballot(in_s) & ballot(in_p)) ? (1u << i) : 0u;
          }
        }
        while (need) {
          const uint32_t t = first + c0 + (uint32_t)__builtin_ctz(need);
          need &= need - 1u;
          R        d2;
          uint32_t f = 0;
          if constexpr (sizeof(R) == 4) {
            const TriRec rec = sload_record(m.tris + t);
            if constexpr (!FACE) {
              if (!tri_eval32_culled(rec, px, py, pz, cull, d2)) continue;
              if (COUNT) w_tris++;
            }
            else {
              // the distance to the triangle's plane is a lower bound on the distance to it
              const float fd = rec.n[0] * (px - rec.v0[0]) + rec.n[1] * (py - rec.v0[1]) + rec.n[2] * (pz - rec.v0[2]);
              if (ballot(fd * fd <= bound) == 0) continue;
              if (COUNT) w_tris++;
              tri_eval32<false>(rec, px, py, pz, d2, f);
              if (d2 <= fc->lim) {
                fc->insert(d2, t);
                if (d2 < best) {
                  best    = d2;
                  btri    = t;
                  fc->lim = fc->band(d2);
                  bound   = fc->lim < bound ? fc->lim : bound;
                }
              }
              continue;
            }
          }
          else {
            if (COUNT) w_tris++;
            const SWords<12> pos = sload<12>(m.tripos + 12 * (size_t)t);
            const SWords<8>  fn  = sload<8>(m.fn64 + 4 * (size_t)t);
            double           fnd[4];
            __builtin_memcpy(fnd, fn.w, sizeof(fnd));
            float posf[12];
            __builtin_memcpy(posf, pos.w, sizeof(posf));
            tri_eval64(posf, fnd, px, py, pz, d2, f);
          }
          if (d2 < best) {
            best  = d2;
            btri  = t;
            bfeat = f;
            bound = d2 < bound ? d2 : bound; // bound stays min(bound, best)
            if constexpr (sizeof(R) == 4 && !FACE) cull = cull_band(bound, m.hmax);
          }
        }
      }
      if (ballot(bound < bound0)) stamp++;
    }
  }
  // FP32: classify the winning triangle's closest feature (per-lane record load)
  if constexpr (sizeof(R) == 4) {
    if (btri != kInvalid) {
      float d2w;
      tri_eval32<true>(m.tris[btri], px, py, pz, d2w, bfeat);
    }
  }
}

template <int K, int NB, class R, bool COUNT, bool WMIN = (sizeof(R) == 8), bool FACE = false, bool PAIRS = false>
__device__ __forceinline__ void traverse(const MeshDev& m, R px, R py, R pz, R bound, R& best, uint32_t& btri,
                                         uint32_t& bfeat, uint32_t& n_inner, uint32_t& n_tris, uint32_t& w_inner, uint32_t& w_tris,
                                         FaceCands* fc = nullptr, PairLds* pl = nullptr, PairCounts* pc = nullptr)
{
  const Entry root{0u, 0u, m.root_lo[0], m.root_lo[1], m.root_lo[2], m.root_hi[0], m.root_hi[1], m.root_hi[2]};
  traverse_from<K, NB, R, COUNT, WMIN, FACE, PAIRS>(m, root, px, py, pz, bound, best, btri, bfeat, n_inner, n_tris,
                                                    w_inner, w_tris, fc, pl, pc);
}

} // namespace dev
} // namespace ebgpu
