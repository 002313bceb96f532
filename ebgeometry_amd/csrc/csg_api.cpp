// C-ABI for CSG programs: validate the instruction list, build one 4-wide BVH (host SAH over
// the primitives' bounding volumes) per BVH union, upload, query.
//
// Query pipeline (caller's stream, no host synchronisation): Morton presort of the queries ->
// FP32 interpreter pass -> FP64 interpreter pass over the queries with |v| < recheck_eps * the
// program's length scale.
#include "host_common.hpp"
#include "mesh_build.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <deque>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

static_assert(int(EBG_CSG_MESH) == int(ebgpu::kCsgMesh) && int(EBG_CSG_POP_POINT) == int(ebgpu::kCsgPopPoint) &&
                int(EBG_CSG_PUSH_CONST) == int(ebgpu::kCsgPushConst) &&
                int(EBG_CSG_BVH_SMOOTH_UNION) == int(ebgpu::kCsgBvhSmoothUnion) &&
                int(EBG_CSG_BVH_SMOOTH_UNION) + 1 == int(ebgpu::kCsgNumOps),
              "public and device opcodes agree");

using ebgpu::DevBuf;
using ebgpu::fail;
using ebgpu::upload;
using ebgpu::Workspace;

struct ebg_csg
{
  std::vector<ebg_mesh*> meshes; // retained
  DevBuf                 ops, ops64, bvhs, prims, meshdev, status;
  std::deque<DevBuf>     nodes;  // one node array per BVH union (a deque: DevBuf does not move)
  ebgpu::CsgDev          dev{};
  int                    KM = 4, banks = 1;
  double                 scale = 1.0;
  std::mutex             mu;
  Workspace              ws;
  std::mutex             grid_mu;
  DevBuf                 gridpts;

  ~ebg_csg()
  {
    for (ebg_mesh* m : meshes) ebgpu::mesh_release(m);
  }
};

namespace {

float round_down(double x)
{
  float f = float(x);
  if (double(f) > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
  return f;
}
float round_up(double x)
{
  float f = float(x);
  if (double(f) < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
  return f;
}

// Stack effect of an instruction: values needed on the stack, values after minus before, and
// saved points needed / change.
struct Effect
{
  int need, dv, needp, dp;
};

bool effect_of(uint32_t op, Effect& e)
{
  using namespace ebgpu;
  if (op <= kCsgMesh || op == kCsgPushConst || op == kCsgBvhUnion || op == kCsgBvhSmoothUnion) e = {0, 1, 0, 0};
  else if (op >= kCsgTranslate && op <= kCsgRepeat) e = {0, 0, 0, 1};
  else if (op == kCsgPopPoint) e = {0, 0, 1, -1};
  else if (op >= kCsgScaleValue && op <= kCsgAnnular) e = {1, 0, 0, 0};
  else if (op == kCsgMinAcc || op == kCsgMaxAcc || op == kCsgAddAcc || op == kCsgSmoothMinEnd ||
           op == kCsgSmoothMaxEnd || op == kCsgDifference)
    e = {2, -1, 0, 0};
  else if (op == kCsgTwoMinAcc || op == kCsgTwoMaxAcc) e = {3, -1, 0, 0};
  else return false;
  return true;
}

// Walk ops[i0, i1) from depths (v, p), checking every instruction; BVH unions recurse into their
// primitives' subprograms (which may hold no BVH union). Tracks the largest depths reached.
bool simulate(const ebg_csg_op* ops, uint64_t num_ops, const ebg_csg_prim* prims, uint64_t num_prims,
              uint64_t num_meshes, uint64_t i0, uint64_t i1, bool allow_bvh, int& v, int& p, int& vmax, int& pmax,
              std::string& err)
{
  using namespace ebgpu;
  for (uint64_t i = i0; i < i1; i++) {
    const ebg_csg_op& o   = ops[i];
    const std::string at  = "instruction " + std::to_string(i) + ": ";
    Effect            e{};
    if (o.op < 0 || !effect_of(uint32_t(o.op), e)) return err = at + "unknown opcode", false;
    for (int k = 0; k < 12; k++)
      if (!std::isfinite(o.p[k]) && !(o.op == kCsgPushConst && k == 0))
        return err = at + "non-finite parameter", false;
    if (v < e.need) return err = at + "needs more values on the stack", false;
    if (p < e.needp) return err = at + "no saved point to restore", false;
    if (o.op == kCsgMesh && (o.n < 0 || uint64_t(o.n) >= num_meshes)) return err = at + "mesh index out of range", false;
    if ((o.op == kCsgRotate || o.op == kCsgInfCylinder) && (o.n < 0 || o.n > 2))
      return err = at + "axis must be 0, 1 or 2", false;
    if (o.op == kCsgScalePoint && !(o.p[0] != 0.0)) return err = at + "zero scale", false;
    if (o.op == kCsgRepeat)
      for (int c = 0; c < 3; c++)
        if (!(o.p[c] != 0.0)) return err = at + "zero repetition period", false;
    if ((o.op == kCsgSmoothMinEnd || o.op == kCsgBvhSmoothUnion) && (o.n < 0 || o.n > 1))
      return err = at + "smooth kind must be 0 (quadratic) or 1 (exponential)", false;
    if (o.op == kCsgBvhUnion || o.op == kCsgBvhSmoothUnion) {
      if (!allow_bvh) return err = at + "a BVH union inside a BVH union's primitive", false;
      if (o.a < 0 || o.b < 1 || uint64_t(o.a) + uint64_t(o.b) > num_prims)
        return err = at + "primitive range out of bounds", false;
      for (int64_t j = o.a; j < int64_t(o.a) + o.b; j++) {
        const ebg_csg_prim& pr = prims[j];
        if (pr.count < 1 || uint64_t(pr.start) + pr.count > num_ops)
          return err = at + "primitive " + std::to_string(j) + ": subprogram out of bounds", false;
        for (int c = 0; c < 3; c++)
          if (!std::isfinite(pr.bv_lo[c]) || !std::isfinite(pr.bv_hi[c]) || !(pr.bv_lo[c] <= pr.bv_hi[c]))
            return err = at + "primitive " + std::to_string(j) + ": bad bounding volume", false;
        int sv = v, sp = p;
        if (!simulate(ops, num_ops, prims, num_prims, num_meshes, pr.start, uint64_t(pr.start) + pr.count, false, sv,
                      sp, vmax, pmax, err))
          return false;
        if (sv != v + 1 || sp != p)
          return err = at + "primitive " + std::to_string(j) + ": subprogram must leave one value and no point", false;
      }
    }
    v += e.dv;
    p += e.dp;
    vmax = std::max(vmax, v);
    pmax = std::max(pmax, p);
  }
  return true;
}

} // namespace

extern "C" {

ebg_status ebg_csg_create(const ebg_csg_op* ops, uint64_t num_ops, uint64_t main_count, const ebg_csg_prim* prims,
                          uint64_t num_prims, ebg_mesh* const* meshes, uint64_t num_meshes, double length_scale,
                          ebg_csg** out)
{
  using namespace ebgpu;
  if (!out) return fail(EBG_INVALID_ARG, "out is null");
  *out = nullptr;
  if (!ops || num_ops == 0 || main_count == 0 || main_count > num_ops)
    return fail(EBG_INVALID_ARG, "a program needs 1 <= main_count <= num_ops instructions");
  if (num_ops > 0x7FFFFFFFull || num_prims > 0x7FFFFFFFull) return fail(EBG_INVALID_ARG, "program too large");
  if (num_prims && !prims) return fail(EBG_INVALID_ARG, "prims is null");
  if (num_meshes && !meshes) return fail(EBG_INVALID_ARG, "meshes is null");
  if (!(length_scale >= 0.0) || !std::isfinite(length_scale)) return fail(EBG_INVALID_ARG, "bad length scale");
  int KM = 0, mesh_banks = 1;
  for (uint64_t j = 0; j < num_meshes; j++) {
    if (!meshes[j]) return fail(EBG_INVALID_ARG, "null mesh handle");
    if (KM && meshes[j]->K != KM) return fail(EBG_INVALID_ARG, "all program meshes must have the same branching");
    KM         = meshes[j]->K;
    mesh_banks = std::max(mesh_banks, meshes[j]->banks);
  }
  if (!KM) KM = 4;

  int         v = 0, p = 0, vmax = 0, pmax = 0;
  std::string err;
  if (!simulate(ops, num_ops, prims, num_prims, num_meshes, 0, main_count, true, v, p, vmax, pmax, err))
    return fail(EBG_INVALID_ARG, err);
  if (v != 1 || p != 0) return fail(EBG_INVALID_ARG, "the program must leave one value and no saved point");
  if (vmax > kCsgMaxValues)
    return fail(EBG_UNSUPPORTED, "program needs " + std::to_string(vmax) + " stack values (max " +
                                   std::to_string(kCsgMaxValues) + ")");
  if (pmax > kCsgMaxPoints)
    return fail(EBG_UNSUPPORTED, "program needs " + std::to_string(pmax) + " saved points (max " +
                                   std::to_string(kCsgMaxPoints) + ")");

  // device records; BVH unions get their own tree over their primitives' boxes
  std::vector<CsgInstr>   rec(num_ops);
  std::vector<CsgInstr64> rec64(num_ops);
  for (uint64_t i = 0; i < num_ops; i++) {
    CsgInstr   r{};
    CsgInstr64 r64{};
    r.op = r64.op = uint32_t(ops[i].op);
    r.n = r64.n = uint32_t(ops[i].n);
    r.a = r64.a = uint32_t(ops[i].a);
    r.b = r64.b = uint32_t(ops[i].b);
    for (int k = 0; k < 12; k++) {
      r.p[k]   = float(ops[i].p[k]);
      r64.p[k] = ops[i].p[k];
    }
    // smoothing lengths divide: keep them above zero in FP32 too (the reference clamps to the
    // smallest normal number of its precision, CSGImplem.hpp:383)
    if (r.op == kCsgSmoothMinEnd || r.op == kCsgSmoothMaxEnd || r.op == kCsgBvhSmoothUnion)
      r.p[0] = std::max(r.p[0], std::numeric_limits<float>::min());
    rec[i]   = r;
    rec64[i] = r64;
  }
  auto* g  = new ebg_csg;
  g->KM    = KM;
  g->scale = length_scale > 0.0 ? length_scale : 1.0;
  std::vector<CsgBvhDev> bvh_dev;
  std::vector<CsgPrim>   prim_all;
  int                    bvh_banks = 1;
  std::vector<std::pair<uint64_t, std::size_t>> fix; // (BVH index, offset of its prims in prim_all)
  for (uint64_t i = 0; i < main_count; i++) {
    if (ops[i].op != kCsgBvhUnion && ops[i].op != kCsgBvhSmoothUnion) continue;
    const uint64_t     first = uint64_t(ops[i].a), cnt = uint64_t(ops[i].b);
    std::vector<float> blo(3 * cnt), bhi(3 * cnt);
    for (uint64_t j = 0; j < cnt; j++)
      for (int c = 0; c < 3; c++) {
        blo[3 * j + c] = round_down(prims[first + j].bv_lo[c]);
        bhi[3 * j + c] = round_up(prims[first + j].bv_hi[c]);
      }
    WideBvh wb;
    if (!build_wide_bvh(blo.data(), bhi.data(), cnt, 4, 4, wb, err)) {
      delete g;
      return fail(EBG_INVALID_ARG, err);
    }
    bvh_banks = std::max(bvh_banks, (wb.stack_need + 63) / 64);
    CsgBvhDev b{};
    for (int c = 0; c < 3; c++) {
      b.root_lo[c] = wb.lo[c];
      b.root_hi[c] = wb.hi[c];
    }
    fix.emplace_back(bvh_dev.size(), prim_all.size());
    for (uint64_t j = 0; j < cnt; j++) {
      const ebg_csg_prim& src = prims[first + wb.order[j]];
      CsgPrim             pr{};
      pr.start = src.start;
      pr.count = src.count;
      for (int c = 0; c < 3; c++) {
        pr.bv[c]     = blo[3 * wb.order[j] + c];
        pr.bv[3 + c] = bhi[3 * wb.order[j] + c];
      }
      prim_all.push_back(pr);
    }
    rec[i].a = rec64[i].a = uint32_t(bvh_dev.size());
    bvh_dev.push_back(b);
    g->nodes.emplace_back();
    hipError_t e = upload(g->nodes.back(), wb.nodes);
    if (e != hipSuccess) {
      delete g;
      return fail(e == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, std::string("BVH upload: ") + hipGetErrorString(e));
    }
  }
  int banks = std::max(bvh_banks, mesh_banks);
  if (banks > 4) {
    delete g;
    return fail(EBG_STACK_OVERFLOW, "a BVH union needs a traversal stack of more than 256 entries");
  }
  if (banks == 3) banks = 4;
  g->banks = banks;
  for (uint64_t j = 0; j < num_meshes; j++) {
    mesh_retain(meshes[j]);
    g->meshes.push_back(meshes[j]);
  }
  std::vector<MeshDev> md(std::max<uint64_t>(num_meshes, 1));
  for (uint64_t j = 0; j < num_meshes; j++) md[j] = meshes[j]->dev;
  if (prim_all.empty()) prim_all.resize(1);
  if (bvh_dev.empty()) bvh_dev.resize(1);
  hipError_t e;
  if ((e = upload(g->ops, rec)) != hipSuccess || (e = upload(g->ops64, rec64)) != hipSuccess ||
      (e = upload(g->prims, prim_all)) != hipSuccess || (e = upload(g->meshdev, md)) != hipSuccess ||
      (e = upload(g->status, std::vector<uint32_t>(16, 0u))) != hipSuccess) {
    delete g;
    return fail(e == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, std::string("program upload: ") + hipGetErrorString(e));
  }
  for (const auto& [k, off] : fix) {
    bvh_dev[k].nodes = g->nodes[k].p;
    bvh_dev[k].prims = g->prims.as<CsgPrim>() + off;
  }
  if ((e = upload(g->bvhs, bvh_dev)) != hipSuccess) {
    delete g;
    return fail(e == hipErrorOutOfMemory ? EBG_OOM : EBG_HIP_ERROR, std::string("program upload: ") + hipGetErrorString(e));
  }
  g->dev.ops        = g->ops.as<CsgInstr>();
  g->dev.ops64      = g->ops64.as<CsgInstr64>();
  g->dev.bvhs       = g->bvhs.as<CsgBvhDev>();
  g->dev.meshes     = g->meshdev.as<MeshDev>();
  g->dev.main_count = uint32_t(main_count);
  g->dev.status     = g->status.as<uint32_t>();
  *out              = g;
  return EBG_OK;
}

ebg_status ebg_csg_destroy(ebg_csg* g)
{
  delete g;
  return EBG_OK;
}

ebg_status ebg_csg_query(const ebg_csg* cg, const float* d_pts, uint64_t n, float* d_value, const ebg_query_opts* opts)
{
  if (!cg) return fail(EBG_INVALID_ARG, "program is null");
  if (n == 0) return EBG_OK;
  if (!d_pts || !d_value) return fail(EBG_INVALID_ARG, "null point or output array");
  if (n > 0xFFFFFFFFull) return fail(EBG_INVALID_ARG, "at most 2^32-1 queries per call");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  if (o.learned_bound || o.seed_radius || o.work)
    return fail(EBG_INVALID_ARG, "programs take no learned or seeded bound and no work counters");
  ebg_csg*    g = const_cast<ebg_csg*>(cg);
  hipStream_t s = static_cast<hipStream_t>(o.stream);

  std::lock_guard<std::mutex> lock(g->mu);
  Workspace&                  ws = g->ws;
  EBG_HIP(ws.ensure(n));
  uint32_t* cnt = ws.counters();
  EBG_HIP(ebgpu::launch_fill_u32(cnt, 2, 0u, s));
  const uint32_t* idx = nullptr;
  if (o.presort && n > 1) {
    EBG_HIP(ebgpu::presort(ws.view(), d_pts, n, s));
    idx = ws.perm.as<uint32_t>();
  }
  const bool       recheck = o.fp64_recheck || o.fp64_all;
  ebgpu::QueryArgs a{};
  a.pts           = d_pts;
  a.idx           = idx;
  a.n             = n;
  a.out_d         = d_value;
  a.recheck_abs   = o.fp64_all ? INFINITY : (recheck ? float(o.recheck_eps * g->scale) : 0.0f);
  a.recheck_list  = recheck ? ws.recheck_list() : nullptr;
  a.recheck_count = cnt;
  EBG_HIP(ebgpu::launch_csg32(g->dev, g->KM, g->banks, a, s));
  if (recheck) {
    ebgpu::QueryArgs r = a;
    r.idx              = ws.recheck_list();
    r.count_ptr        = cnt;
    r.recheck_list     = nullptr;
    EBG_HIP(ebgpu::launch_csg64(g->dev, g->KM, g->banks, r, s));
  }
  return EBG_OK;
}

ebg_status ebg_csg_query_host(const ebg_csg* g, const float* pts, uint64_t n, float* value, const ebg_query_opts* opts)
{
  if (!g) return fail(EBG_INVALID_ARG, "program is null");
  if (n == 0) return EBG_OK;
  if (!pts || !value) return fail(EBG_INVALID_ARG, "null point or output array");
  DevBuf p, d;
  EBG_HIP(p.reserve(12 * n));
  EBG_HIP(d.reserve(4 * n));
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  EBG_HIP(hipMemcpyAsync(p.p, pts, 12 * n, hipMemcpyHostToDevice, s));
  const ebg_status st = ebg_csg_query(g, p.as<float>(), n, d.as<float>(), &o);
  if (st != EBG_OK) return st;
  EBG_HIP(hipMemcpyAsync(value, d.p, 4 * n, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_csg_query_grid(const ebg_csg* cg, const double* origin, const double* spacing, const uint64_t* dims,
                              float* d_value, const ebg_query_opts* opts)
{
  if (!cg) return fail(EBG_INVALID_ARG, "program is null");
  ebgpu::GridSpec  gs{};
  uint64_t         n  = 0;
  const ebg_status st = ebgpu::grid_spec(origin, spacing, dims, gs, n);
  if (st != EBG_OK) return st;
  if (n == 0) return EBG_OK;
  if (!d_value) return fail(EBG_INVALID_ARG, "null output array");
  ebg_csg*       g = const_cast<ebg_csg*>(cg);
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t                 s = static_cast<hipStream_t>(o.stream);
  std::lock_guard<std::mutex> lock(g->grid_mu);
  EBG_HIP(g->gridpts.reserve(12 * n));
  EBG_HIP(ebgpu::launch_grid_points(gs, g->gridpts.as<float>(), s));
  return ebg_csg_query(g, g->gridpts.as<float>(), n, d_value, &o);
}

ebg_status ebg_csg_query_grid_host(const ebg_csg* g, const double* origin, const double* spacing,
                                   const uint64_t* dims, float* value, const ebg_query_opts* opts)
{
  if (!g) return fail(EBG_INVALID_ARG, "program is null");
  ebgpu::GridSpec  gs{};
  uint64_t         n  = 0;
  const ebg_status st = ebgpu::grid_spec(origin, spacing, dims, gs, n);
  if (st != EBG_OK) return st;
  if (n == 0) return EBG_OK;
  if (!value) return fail(EBG_INVALID_ARG, "null output array");
  ebg_query_opts o;
  ebg_query_opts_default(&o);
  if (opts) o = *opts;
  hipStream_t s = static_cast<hipStream_t>(o.stream);
  DevBuf      d;
  EBG_HIP(d.reserve(4 * n));
  const ebg_status q = ebg_csg_query_grid(g, origin, spacing, dims, d.as<float>(), &o);
  if (q != EBG_OK) return q;
  EBG_HIP(hipMemcpyAsync(value, d.p, 4 * n, hipMemcpyDeviceToHost, s));
  EBG_HIP(hipStreamSynchronize(s));
  return EBG_OK;
}

ebg_status ebg_csg_check(const ebg_csg* g)
{
  if (!g) return fail(EBG_INVALID_ARG, "program is null");
  uint32_t w = 0;
  EBG_HIP(hipDeviceSynchronize());
  EBG_HIP(hipMemcpy(&w, g->status.p, 4, hipMemcpyDeviceToHost));
  if (w & ebgpu::kStatusStackOverflow)
    return fail(EBG_STACK_OVERFLOW, "a BVH walk of this program needed more stack than its build-time worst case");
  return EBG_OK;
}

} // extern "C"
