// C-ABI for batches of meshes: build (optionally autotuned), plan (packet schedule, buffers,
// workgroup tuning, hipGraph capture), launch.
#include "host_common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

using ebgpu::DevBuf;
using ebgpu::fail;
using ebgpu::upload;

struct ebg_batch
{
  std::vector<ebg_mesh*> meshes; // owned
  DevBuf                 meshdev;
  int                    K = 4, leaf = 8, banks = 1;
  bool                   tuned = false;
  double                 tune_ms[4]{-1, -1, -1, -1};
  double                 build_seconds = 0;
  uint64_t               num_triangles = 0;

  ~ebg_batch()
  {
    for (ebg_mesh* m : meshes) ebgpu::mesh_release(m);
  }
};

struct ebg_batch_plan
{
  ebg_batch*      b = nullptr;
  uint64_t        total = 0;
  uint32_t        num_items = 0, num_packets = 0;
  DevBuf          items, packets, keys, keys_sorted, sorted, scratch, rmask, recheck, counter;
  DevBuf          grids, flags, early; // early packets: per-mesh grid records, packet flags, list
  uint32_t        early_cap = 0;       // 0: no early packets
  int             wg = 256;
  bool            persistent = false;  // one grid that fits the device, wavefronts take packets from a queue
  uint32_t        persist_blocks = 0;
  bool            presort = true, fp64 = true;
  unsigned long long* work = nullptr;
  unsigned long long* pair_work = nullptr;
  hipGraph_t      graph = nullptr;
  hipGraphExec_t  exec  = nullptr;
  hipEvent_t      ev0 = nullptr, ev1 = nullptr;
  bool            timed = false;

  ~ebg_batch_plan()
  {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
  }
};

namespace {

// Deterministic synthetic queries for tuning: half uniform in the 1.2x bounding box, half near
// the surface (area-weighted triangle, uniform barycentric point, offset N(0, 0.02 diag)).
struct Rng
{
  uint64_t s;
  uint64_t next()
  {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal()
  {
    const double u = std::max(uniform(), 1e-300), v = uniform();
    return std::sqrt(-2.0 * std::log(u)) * std::cos(6.283185307179586 * v);
  }
};

std::vector<float> tuning_queries(const float* xyz, uint64_t nv, const uint32_t* tri, uint64_t nt, uint64_t n,
                                  uint64_t seed)
{
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (uint64_t v = 0; v < nv; v++)
    for (int c = 0; c < 3; c++) {
      lo[c] = std::min(lo[c], double(xyz[3 * v + c]));
      hi[c] = std::max(hi[c], double(xyz[3 * v + c]));
    }
  const double        diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                       (hi[2] - lo[2]) * (hi[2] - lo[2]));
  std::vector<double> cdf(nt);
  double              acc = 0;
  for (uint64_t f = 0; f < nt; f++) {
    const float* a = xyz + 3 * tri[3 * f];
    const float* b = xyz + 3 * tri[3 * f + 1];
    const float* c = xyz + 3 * tri[3 * f + 2];
    const double ux = b[0] - a[0], uy = b[1] - a[1], uz = b[2] - a[2];
    const double wx = c[0] - a[0], wy = c[1] - a[1], wz = c[2] - a[2];
    const double cx = uy * wz - uz * wy, cy = uz * wx - ux * wz, cz = ux * wy - uy * wx;
    acc += std::sqrt(cx * cx + cy * cy + cz * cz);
    cdf[f] = acc;
  }
  Rng                rng{seed};
  std::vector<float> q(3 * n);
  for (uint64_t i = 0; i < n; i++) {
    if (i % 2 == 0) {
      for (int c = 0; c < 3; c++) {
        const double mid = 0.5 * (lo[c] + hi[c]), h = 0.6 * (hi[c] - lo[c]);
        q[3 * i + c]     = float(mid - h + 2.0 * h * rng.uniform());
      }
    }
    else {
      const uint64_t f = std::min<uint64_t>(nt - 1, std::lower_bound(cdf.begin(), cdf.end(), rng.uniform() * acc) -
                                                        cdf.begin());
      double         r1 = rng.uniform(), r2 = rng.uniform();
      if (r1 + r2 > 1.0) {
        r1 = 1.0 - r1;
        r2 = 1.0 - r2;
      }
      const float* a = xyz + 3 * tri[3 * f];
      const float* b = xyz + 3 * tri[3 * f + 1];
      const float* c = xyz + 3 * tri[3 * f + 2];
      double       p[3], u[3], w[3];
      for (int k = 0; k < 3; k++) {
        u[k] = b[k] - a[k];
        w[k] = c[k] - a[k];
        p[k] = a[k] + r1 * u[k] + r2 * w[k];
      }
      double nrm[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      const double l = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      const double o = 0.02 * diag * rng.normal() / (l > 0 ? l : 1.0);
      for (int k = 0; k < 3; k++) q[3 * i + k] = float(p[k] + o * nrm[k]);
    }
  }
  return q;
}

// Build meshes[i] for the listed indices on `threads` host threads.
ebg_status build_meshes(const float* const* xyz, const uint64_t* nv, const uint32_t* const* tri, const uint64_t* nt,
                        const std::vector<uint64_t>& which, int K, int leaf, int threads,
                        std::vector<ebg_mesh*>& out)
{
  out.assign(which.size(), nullptr);
  std::vector<ebg_status>  st(which.size(), EBG_OK);
  std::vector<std::string> msg(which.size());
  std::atomic<uint64_t>    next{0};
  auto                     work = [&]() {
    for (;;) {
      const uint64_t k = next.fetch_add(1);
      if (k >= which.size()) return;
      const uint64_t i = which[k];
      ebg_build_opts o{EBG_BUILD_SAH, K, leaf, 0};
      st[k] = ebg_mesh_create(xyz[i], nv[i], tri[i], nt[i], &o, &out[k]);
      if (st[k] != EBG_OK) msg[k] = "mesh " + std::to_string(i) + ": " + ebg_last_error();
    }
  };
  const int                nth = std::max(1, std::min<int>(threads, int(which.size())));
  std::vector<std::thread> pool;
  for (int t = 1; t < nth; t++) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  for (uint64_t k = 0; k < which.size(); k++)
    if (st[k] != EBG_OK) {
      for (ebg_mesh* m : out)
        if (m) ebgpu::mesh_release(m);
      out.clear();
      return fail(st[k], msg[k]);
    }
  return EBG_OK;
}

ebg_status enqueue(ebg_batch_plan* p, hipStream_t s, bool events)
{
  ebg_batch*       b = p->b;
  ebgpu::BatchArgs a{};
  a.items         = p->items.as<ebgpu::BatchItemDev>();
  a.packets       = p->packets.as<uint32_t>();
  a.num_packets   = p->num_packets;
  a.meshes        = b->meshdev.as<ebgpu::MeshDev>();
  a.sorted        = nullptr;
  a.recheck_mask  = p->fp64 ? p->rmask.as<uint64_t>() : nullptr;
  a.recheck_list  = p->fp64 ? p->recheck.as<uint32_t>() : nullptr;
  a.recheck_count = p->counter.as<uint32_t>();
  a.work          = p->work;
  a.pair_work     = p->work ? p->pair_work : nullptr;
  // counters: [0] recheck list, [1] early list, [2] persistent queue (a fill kernel, so that a
  // captured plan clears them on every replay)
  if (p->fp64 || p->early_cap || p->persistent) EBG_HIP(ebgpu::launch_fill_u32(a.recheck_count, 3, 0u, s));
  if (p->presort && p->total > 1) {
    int item_bits = 1;
    while ((1ull << item_bits) < p->num_items) item_bits++;
    const bool narrow = 24 + item_bits <= 32; // 32-bit keys: half the key traffic per pass
    EBG_HIP(ebgpu::launch_batch_keys(a.items, p->num_items, p->total, p->keys.p, narrow, s));
    if (narrow)
      EBG_HIP(ebgpu::radix_sort_pairs32(p->keys.as<uint32_t>(), nullptr, p->keys_sorted.as<uint32_t>(),
                                        p->sorted.as<uint32_t>(), p->total, 24 + item_bits, p->scratch.p, false, s));
    else
      EBG_HIP(ebgpu::radix_sort_pairs(p->keys.as<uint64_t>(), nullptr, p->keys_sorted.as<uint64_t>(),
                                      p->sorted.as<uint32_t>(), p->total, 24 + item_bits, p->scratch.p, s));
    a.sorted = p->sorted.as<uint32_t>();
  }
  if (p->early_cap) {
    a.grids        = p->grids.as<ebgpu::GridDev>();
    a.flag         = p->flags.as<uint8_t>();
    a.early_list   = p->early.as<uint32_t>();
    a.early_count  = p->counter.as<uint32_t>() + 1;
    a.early_cap    = p->early_cap;
    a.early_blocks = (p->early_cap + uint32_t(p->wg / 64) - 1) / uint32_t(p->wg / 64);
    EBG_HIP(ebgpu::launch_batch_mark(a, s));
  }
  if (p->persistent) {
    a.queue          = p->counter.as<uint32_t>() + 2;
    a.queue_total    = (a.early_blocks ? a.early_cap : 0u) + p->num_packets;
    a.persist_blocks = p->persist_blocks;
  }
  if (events) EBG_HIP(hipEventRecord(p->ev0, s));
  EBG_HIP(ebgpu::launch_batch32(a, b->K, b->banks, p->wg, p->work != nullptr, s));
  if (events) EBG_HIP(hipEventRecord(p->ev1, s));
  if (p->fp64) {
    // the list follows the packet table (presort order), so FP64 wavefronts hold neighbouring
    // queries
    EBG_HIP(ebgpu::launch_recheck_compact(a.recheck_mask, uint64_t(p->num_packets) * 64, nullptr, a.recheck_list,
                                          a.recheck_count, s));
    EBG_HIP(ebgpu::launch_batch64(a, b->K, b->banks, s));
  }
  return EBG_OK;
}

// Milliseconds between two recorded events, or -1; a failure is cleared from the runtime's
// last-error state so that it does not surface in the caller's next HIP call.
// Early packets of a plan: meshes from 16,384 faces get a cost grid; packets at >= 2x their
// mesh's grid mean start first (256-mesh workload: 66.8 -> 55.4 ms; 1.5x: 56.3, 3x: 55.4).
constexpr uint64_t kEarlyMinFaces = 16384;
constexpr double   kEarlyFactor   = 2.0;

double elapsed_ms(hipEvent_t a, hipEvent_t b)
{
  float ms = -1.0f;
  if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return -1.0;
  }
  return ms;
}

} // namespace

extern "C" {

ebg_status ebg_batch_create(const float* const* xyz, const uint64_t* nv, const uint32_t* const* tri,
                            const uint64_t* nt, uint64_t num_meshes, const ebg_batch_opts* opts, ebg_batch** out)
{
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (num_meshes == 0 || !xyz || !nv || !tri || !nt) return fail(EBG_INVALID_ARG, "empty batch");
  if (num_meshes > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "too many meshes");
  for (uint64_t i = 0; i < num_meshes; i++) {
    if (!xyz[i] || !tri[i] || nv[i] < 3 || nt[i] < 1)
      return fail(EBG_INVALID_ARG, "mesh " + std::to_string(i) + ": empty mesh");
    for (uint64_t k = 0; k < 3 * nt[i]; k++)
      if (tri[i][k] >= nv[i]) return fail(EBG_INVALID_ARG, "mesh " + std::to_string(i) + ": index out of range");
  }
  ebg_batch_opts o{};
  if (opts) o = *opts;
  const int threads = o.threads > 0 ? o.threads : std::max(1, int(std::thread::hardware_concurrency()));
  int       K = o.branching ? o.branching : 4, leaf = o.leaf_size ? o.leaf_size : 8;
  if (K != 4 && K != 8) return fail(EBG_INVALID_ARG, "branching must be 4 or 8");
  if (leaf < 1 || leaf > ebgpu::kMaxLeaf) return fail(EBG_INVALID_ARG, "leaf_size must be 1..16");

  const auto t0 = std::chrono::steady_clock::now();
  auto*      b  = new ebg_batch;
  if (o.autotune) {
    // the two largest meshes, 2^17 synthetic queries each, plain traversal time per candidate
    std::vector<uint64_t> order(num_meshes);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) { return nt[x] > nt[y]; });
    order.resize(std::min<uint64_t>(2, num_meshes));
    const uint64_t                  nq = 1u << 17;
    std::vector<std::vector<float>> qh;
    for (uint64_t i : order) qh.push_back(tuning_queries(xyz[i], nv[i], tri[i], nt[i], nq, 1234 + i));
    DevBuf qd, dd;
    hipError_t e;
    if ((e = qd.reserve(12 * nq * order.size())) != hipSuccess || (e = dd.reserve(4 * nq)) != hipSuccess) {
      delete b;
      return fail(EBG_OOM, std::string("tuning buffers: ") + hipGetErrorString(e));
    }
    for (std::size_t k = 0; k < order.size(); k++)
      (void)hipMemcpy(qd.as<float>() + 3 * nq * k, qh[k].data(), 12 * nq, hipMemcpyHostToDevice);
    const int cand[4][2] = {{4, 8}, {4, 16}, {8, 8}, {8, 16}};
    double    best       = 1e300;
    for (int c = 0; c < 4; c++) {
      std::vector<ebg_mesh*> ms;
      const ebg_status       st = build_meshes(xyz, nv, tri, nt, order, cand[c][0], cand[c][1], threads, ms);
      if (st == EBG_STACK_OVERFLOW) continue;
      if (st != EBG_OK) {
        delete b;
        return st;
      }
      double total = 0;
      for (std::size_t k = 0; k < ms.size(); k++) {
        ebg_query_opts qo;
        ebg_query_opts_default(&qo);
        qo.fp64_recheck = 0;
        double tmin     = 1e300;
        for (int rep = 0; rep < 3; rep++) {
          const ebg_status qs = ebg_query_sdf(ms[k], qd.as<float>() + 3 * nq * k, nq, dd.as<float>(), nullptr, &qo);
          if (qs != EBG_OK) {
            for (ebg_mesh* m : ms) ebgpu::mesh_release(m);
            delete b;
            return qs;
          }
          tmin = std::min(tmin, ebg_mesh_last_traversal_ms(ms[k]));
        }
        total += tmin;
      }
      for (ebg_mesh* m : ms) ebgpu::mesh_release(m);
      b->tune_ms[c] = total;
      if (total < best) {
        best = total;
        K    = cand[c][0];
        leaf = cand[c][1];
      }
    }
    b->tuned = true;
  }
  std::vector<uint64_t> all(num_meshes);
  std::iota(all.begin(), all.end(), 0);
  const ebg_status st = build_meshes(xyz, nv, tri, nt, all, K, leaf, threads, b->meshes);
  if (st != EBG_OK) {
    delete b;
    return st;
  }
  b->K    = K;
  b->leaf = leaf;
  std::vector<ebgpu::MeshDev> md(num_meshes);
  for (uint64_t i = 0; i < num_meshes; i++) {
    md[i]    = b->meshes[i]->dev;
    b->banks = std::max(b->banks, b->meshes[i]->banks);
    b->num_triangles += b->meshes[i]->nt;
  }
  const hipError_t e = upload(b->meshdev, md);
  if (e != hipSuccess) {
    delete b;
    return fail(EBG_HIP_ERROR, std::string("batch upload: ") + hipGetErrorString(e));
  }
  b->build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *out             = b;
  return EBG_OK;
}

ebg_status ebg_batch_destroy(ebg_batch* batch)
{
  delete batch;
  return EBG_OK;
}

ebg_status ebg_batch_get_info(const ebg_batch* b, ebg_batch_info* info)
{
  if (!b || !info) return fail(EBG_INVALID_ARG, "null argument");
  std::memset(info, 0, sizeof(*info));
  info->num_meshes     = b->meshes.size();
  info->num_triangles  = b->num_triangles;
  info->branching      = b->K;
  info->leaf_size      = b->leaf;
  info->stack_capacity = 64 * b->banks;
  info->tuned          = b->tuned ? 1 : 0;
  for (int c = 0; c < 4; c++) info->tune_ms[c] = b->tune_ms[c];
  info->build_seconds = b->build_seconds;
  return EBG_OK;
}

ebg_status ebg_batch_mesh(const ebg_batch* b, uint64_t i, ebg_mesh** mesh)
{
  if (!b || !mesh) return fail(EBG_INVALID_ARG, "null argument");
  if (i >= b->meshes.size()) return fail(EBG_INVALID_ARG, "mesh index out of range");
  *mesh = b->meshes[i];
  return EBG_OK;
}

ebg_status ebg_batch_plan_create(ebg_batch* b, const ebg_batch_item* items, uint64_t num_items,
                                 const ebg_query_opts* opts, int32_t workgroup, int32_t capture,
                                 ebg_batch_plan** out)
{
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (!b) return fail(EBG_INVALID_ARG, "batch is null");
  if (num_items && !items) return fail(EBG_INVALID_ARG, "items is null");
  if (num_items > 0xFFFFFFull) return fail(EBG_INVALID_ARG, "too many items");
  const bool want_persistent = (workgroup & EBG_PLAN_PERSISTENT) != 0;
  workgroup &= ~EBG_PLAN_PERSISTENT;
  if (workgroup != 0 && workgroup != 64 && workgroup != 128 && workgroup != 256)
    return fail(EBG_INVALID_ARG, "workgroup must be 0, 64, 128 or 256 (optionally | EBG_PLAN_PERSISTENT)");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  if (o.learned_bound || o.seed_radius) return fail(EBG_INVALID_ARG, "batches take no learned or seeded bound");

  // items -> device records and the packet schedule (most expensive meshes first)
  std::vector<ebgpu::BatchItemDev> rec(num_items);
  uint64_t                         total = 0;
  for (uint64_t i = 0; i < num_items; i++) {
    const ebg_batch_item& it = items[i];
    if (it.mesh >= b->meshes.size()) return fail(EBG_INVALID_ARG, "item " + std::to_string(i) + ": bad mesh index");
    if (it.n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "item " + std::to_string(i) + ": too many queries");
    if (it.n && (!it.d_pts || !it.d_dist))
      return fail(EBG_INVALID_ARG, "item " + std::to_string(i) + ": null point or output array");
    const ebg_mesh*      m = b->meshes[it.mesh];
    ebgpu::BatchItemDev& r = rec[i];
    std::memset(&r, 0, sizeof(r));
    r.pts  = it.d_pts;
    r.out  = it.d_dist;
    r.tri  = it.d_tri;
    r.base = total;
    r.n    = uint32_t(it.n);
    r.mesh = it.mesh;
    for (int c = 0; c < 3; c++) {
      const double mid = 0.5 * (m->lo[c] + m->hi[c]), h = 0.625 * (m->hi[c] - m->lo[c]);
      r.box_lo[c]      = float(mid - h);
      r.box_inv[c]     = h > 0 ? float(256.0 / (2.0 * h)) : 0.0f;
    }
    r.recheck_abs = o.fp64_all ? INFINITY : float(o.recheck_eps * m->diag);
    r.slack_abs   = float(1e-5 * m->diag);
    total += it.n;
  }
  if (total > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 queries per plan");
  std::vector<uint64_t> ord(num_items);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](uint64_t x, uint64_t y) {
    return b->meshes[items[x].mesh]->nt > b->meshes[items[y].mesh]->nt;
  });
  std::vector<uint32_t> pk;
  for (uint64_t i : ord)
    for (uint64_t f = 0; f < items[i].n; f += 64) {
      pk.push_back(uint32_t(i));
      pk.push_back(uint32_t(f));
      pk.push_back(0u);
      pk.push_back(0u);
    }

  auto* p        = new ebg_batch_plan;
  p->b           = b;
  p->total       = total;
  p->num_items   = uint32_t(num_items);
  p->num_packets = uint32_t(pk.size() / 4);
  p->presort     = o.presort != 0;
  p->fp64        = (o.fp64_recheck || o.fp64_all) && p->num_packets > 0;
  p->work        = reinterpret_cast<unsigned long long*>(o.work);
  p->pair_work   = reinterpret_cast<unsigned long long*>(o.pair_work);
  p->wg          = workgroup ? workgroup : 256;
  p->persistent  = want_persistent;
  auto bail      = [&](hipError_t e, const char* what) {
    delete p;
    return fail(e == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = upload(p->items, rec)) != hipSuccess) return bail(e, "items");
  if ((e = upload(p->packets, pk)) != hipSuccess) return bail(e, "packets");
  const uint64_t cap = std::max<uint64_t>(total, 1);
  if ((e = p->keys.reserve(8 * cap)) != hipSuccess) return bail(e, "keys");
  if ((e = p->keys_sorted.reserve(8 * cap)) != hipSuccess) return bail(e, "keys");
  if ((e = p->sorted.reserve(4 * cap)) != hipSuccess) return bail(e, "sort");
  if ((e = p->scratch.reserve(ebgpu::radix_sort_scratch_bytes(cap))) != hipSuccess) return bail(e, "sort scratch");
  // recheck entries are packet << 6 | lane in 32 bits; at most one per query
  if (p->fp64 && p->num_packets > (1u << 26)) {
    delete p;
    return fail(EBG_INVALID_ARG, "FP64 recheck of more than 2^26 packets in one plan");
  }
  if ((e = p->rmask.reserve(8 * std::max<uint64_t>(p->num_packets, 1))) != hipSuccess) return bail(e, "recheck");
  if ((e = p->recheck.reserve(4 * (p->fp64 ? cap : 1))) != hipSuccess) return bail(e, "recheck");
  if ((e = p->counter.reserve(256)) != hipSuccess) return bail(e, "counter");
  if ((e = hipEventCreate(&p->ev0)) != hipSuccess || (e = hipEventCreate(&p->ev1)) != hipSuccess)
    return bail(e, "events");

  hipStream_t s = static_cast<hipStream_t>(o.stream);
  // Early packets: queries near a mesh's medial axis cost a packet tens of times the median, and
  // one that starts late ends the launch. Meshes of at least kEarlyMinFaces get their packet
  // cost grid (built once per mesh), and packets at >= kEarlyFactor x their mesh's grid mean
  // start first.
  if (p->num_packets > 0) {
    std::vector<ebgpu::GridDev> gd(b->meshes.size());
    std::vector<char>           used(b->meshes.size(), 0);
    for (uint64_t i = 0; i < num_items; i++) used[items[i].mesh] = 1;
    bool any = false;
    for (std::size_t k = 0; k < b->meshes.size(); k++) {
      ebgpu::GridDev& g = gd[k];
      g.threshold       = ebgpu::kInvalid;
      ebg_mesh* m       = b->meshes[k];
      if (!used[k] || m->nt < kEarlyMinFaces) continue;
      {
        std::lock_guard<std::recursive_mutex> lock(m->mu);
        const ebg_status gs = ebgpu::ensure_cost_grid(m, s);
        if (gs != EBG_OK) {
          delete p;
          return gs;
        }
      }
      g.cost = m->cost_grid.as<uint32_t>();
      for (int c = 0; c < 3; c++) {
        g.lo[c]    = m->grid_lo[c];
        g.scale[c] = m->grid_scale[c];
      }
      g.g         = ebgpu::kCostGrid;
      g.threshold = uint32_t(std::max(1.0, kEarlyFactor * m->grid_mean));
      any         = true;
    }
    if (any) {
      p->early_cap = p->num_packets / 8 + 64;
      if ((e = upload(p->grids, gd)) != hipSuccess) return bail(e, "grids");
      if ((e = p->flags.reserve(p->num_packets)) != hipSuccess) return bail(e, "flags");
      if ((e = p->early.reserve(4 * (std::size_t)p->early_cap)) != hipSuccess) return bail(e, "early list");
    }
  }

  auto set_shape = [&](int wg, bool persistent) -> hipError_t {
    p->wg         = wg;
    p->persistent = persistent;
    return persistent ? ebgpu::batch32_persistent_blocks(b->K, b->banks, wg, p->work != nullptr, &p->persist_blocks)
                      : hipSuccess;
  };
  if ((e = set_shape(p->wg, p->persistent)) != hipSuccess) return bail(e, "occupancy");
  if (!workgroup && p->num_packets > 0) {
    // pick the launch shape by timing the traversal on the items as they are: one wavefront per
    // packet in workgroups of 256, 128 or 64 lanes, or the persistent grid (64 or 256 lanes)
    double best = 1e300;
    int    pick = 256;
    bool   pick_p = false;
    const std::pair<int, bool> shapes[] = {{256, false}, {128, false}, {64, false}, {64, true}, {256, true}};
    for (const auto& sh : shapes) {
      if ((e = set_shape(sh.first, sh.second)) != hipSuccess) return bail(e, "occupancy");
      double tmin = 1e300;
      for (int rep = 0; rep < 3; rep++) {
        const ebg_status st = enqueue(p, s, true);
        if (st != EBG_OK) {
          delete p;
          return st;
        }
        tmin = std::min(tmin, elapsed_ms(p->ev0, p->ev1));
      }
      if (tmin < best) {
        best   = tmin;
        pick   = sh.first;
        pick_p = sh.second;
      }
    }
    if ((e = set_shape(pick, pick_p)) != hipSuccess) return bail(e, "occupancy");
  }
  if (capture) {
    hipStream_t cs = nullptr;
    if ((e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking)) != hipSuccess) return bail(e, "capture stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    ebg_status st = e == hipSuccess ? enqueue(p, cs, false) : EBG_HIP_ERROR;
    hipGraph_t g  = nullptr;
    const hipError_t e2 = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) return bail(e, "begin capture");
    if (st != EBG_OK) {
      if (g) (void)hipGraphDestroy(g);
      delete p;
      return st;
    }
    if (e2 != hipSuccess) return bail(e2, "end capture");
    p->graph = g;
    if ((e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0)) != hipSuccess) return bail(e, "instantiate");
  }
  *out = p;
  return EBG_OK;
}

ebg_status ebg_batch_plan_launch(ebg_batch_plan* p, void* stream)
{
  if (!p) return fail(EBG_INVALID_ARG, "plan is null");
  if (p->num_packets == 0) return EBG_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (p->exec) {
    EBG_HIP(hipEventRecord(p->ev0, s));
    EBG_HIP(hipGraphLaunch(p->exec, s));
    EBG_HIP(hipEventRecord(p->ev1, s));
  }
  else {
    const ebg_status st = enqueue(p, s, true);
    if (st != EBG_OK) return st;
  }
  p->timed = true;
  return EBG_OK;
}

ebg_status ebg_batch_plan_destroy(ebg_batch_plan* p)
{
  delete p;
  return EBG_OK;
}

int32_t ebg_batch_plan_workgroup(const ebg_batch_plan* p)
{
  return p ? (p->wg | (p->persistent ? EBG_PLAN_PERSISTENT : 0)) : 0;
}

int64_t ebg_batch_plan_last_early_packets(const ebg_batch_plan* p)
{
  if (!p) return -1;
  if (!p->early_cap || !p->timed) return 0;
  uint32_t c = 0;
  if (hipDeviceSynchronize() != hipSuccess) return -1;
  if (hipMemcpy(&c, p->counter.as<uint32_t>() + 1, 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return std::min<int64_t>(c, p->early_cap);
}

double ebg_batch_plan_last_traversal_ms(const ebg_batch_plan* p)
{
  if (!p || !p->timed) return -1.0;
  return elapsed_ms(p->ev0, p->ev1);
}

ebg_status ebg_batch_query(ebg_batch* b, const ebg_batch_item* items, uint64_t num_items, const ebg_query_opts* opts)
{
  ebg_batch_plan*  p  = nullptr;
  const ebg_status st = ebg_batch_plan_create(b, items, num_items, opts, 256, 0, &p);
  if (st != EBG_OK) return st;
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  ebg_status r = ebg_batch_plan_launch(p, o.stream);
  // the plan's buffers must outlive the enqueued work
  if (r == EBG_OK && hipStreamSynchronize(static_cast<hipStream_t>(o.stream)) != hipSuccess)
    r = fail(EBG_HIP_ERROR, "batch query failed");
  ebg_batch_plan_destroy(p);
  return r;
}

} // extern "C"
