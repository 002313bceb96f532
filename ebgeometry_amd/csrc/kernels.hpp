// Host-side launchers of the HIP kernels (defined in the .hip translation units).
#pragma once

#include "ebgpu_internal.hpp"

#include <hip/hip_runtime.h>

namespace ebgpu {

// Arguments of one traversal launch over a list of queries.
struct QueryArgs
{
  const float*        pts;       // original point array (3 floats per query)
  const uint32_t*     idx;       // query index per thread (presort order or a compacted list); null = identity
  const uint32_t*     count_ptr; // device count of idx entries (list launches); null = use n
  uint64_t            n;         // threads to launch (upper bound when count_ptr is set)
  const float*        seed_r;    // per-query seeded cull radius (indexed by query); null = +inf
  float*              out_d;     // indexed by query
  uint32_t*           out_tri;   // indexed by query; may be null
  float               recheck_abs; // |d| below this goes to the recheck list (0 = off)
  uint32_t*           recheck_list;
  uint32_t*           recheck_count;
  // non-null: instead of appending to recheck_list, each packet pk (threads 64 pk ..) writes the
  // lane mask of its recheck queries here; launch_recheck_compact then builds the list
  uint64_t*           recheck_mask;
  uint32_t*           fallback_list;  // queries whose seeded radius held no triangle
  uint32_t*           fallback_count;
  unsigned long long* work;      // [inner nodes, triangles] accumulated (null = not counted)
  unsigned long long* pair_work; // counting launches: [pairs, pair batches, culling loads, slab loads]
  float               slack_abs; // FP64 recheck: bound = (|d| + slack_abs + slack_rel*|d|)^2
  float               slack_rel;
  uint32_t*           packet_cost; // counting launches: per packet, the nodes + triangles its wavefront fetched
  // Split packets (split_blocks = 0: none). The first split_blocks workgroups of the FP32 launch
  // stride over the items (j, r), j < min(*split_count, split_cap), r < split_nroots: walk subtree
  // split_roots[r] for the 64 queries of packet split_list[j] (unseeded) and fold each lane's
  // (d^2, triangle) into split_keys[64 j + lane] with an atomic minimum. The other workgroups
  // skip the packets flagged in heavy_flag; a resolve launch finishes the split packets.
  const uint8_t*      heavy_flag;
  const SplitRoot*    split_roots;
  const uint32_t*     split_list;
  const uint32_t*     split_count;
  unsigned long long* split_keys;
  uint32_t            split_nroots, split_cap, split_blocks;
  // Early packets: the next early_blocks workgroups take early_list[0 .. min(*early_count,
  // early_cap)) (flagged too), so that the costlier unsplit packets start first.
  const uint32_t*     early_list;
  const uint32_t*     early_count;
  uint32_t            early_cap, early_blocks;
};

// Cost flags of the split schedule, set while the presort keys are made: a query whose
// cost-grid value (grid of g^3 cells over [lo, lo + g/scale)) reaches heavy / early sets bit 31
// / bit 30 of its 30-bit presort key. The sort ignores both bits and carries them to the
// sorted keys, so the packet marking reads them in sorted order without gathering points.
struct CostFlags
{
  const uint32_t* grid;
  float           lo[3], scale[3];
  uint32_t        g, heavy, early;
};

// Expensive packets: a packet with a heavy-flagged query is listed (up to cap), flagged, and its
// 64 result keys are reset; otherwise a packet with an early-flagged query goes to the early
// list (up to early_cap). skeys are the sorted presort keys; count must be zero on entry.
struct HeavyMarkArgs
{
  const uint32_t*     skeys;
  uint64_t            n;
  uint32_t            cap;
  uint32_t*           list;
  uint32_t*           count;
  uint8_t*            flag;
  unsigned long long* keys; // cap * 64
  uint32_t            early_cap;
  uint32_t*           early_list;
  uint32_t*           early_count;
};
hipError_t launch_heavy_mark(const HeavyMarkArgs& h, hipStream_t s);

// Recheck list from the per-packet masks of n threads: one wavefront per 64 consecutive packets
// appends their flagged queries (idx[thread], or the thread) with one atomic on *count, in packet
// order, so each run of the list covers a compact stretch of the presort curve.
hipError_t launch_recheck_compact(const uint64_t* mask, uint64_t n, const uint32_t* idx, uint32_t* list,
                                  uint32_t* count, hipStream_t s);

// After the FP32 launch: the split packets' winning triangles classified, signed and written like
// the plain pass (outputs, recheck list).
hipError_t launch_sdf32_resolve(const MeshDev& m, const QueryArgs& a, hipStream_t s);

// FP32 packet traversal. K = 4 or 8; banks = stack capacity / 64 (1..4).
hipError_t launch_sdf32(const MeshDev& m, int K, int banks, const QueryArgs& a, bool count, hipStream_t s);
// FP64 traversal that recomputes distance, closest face and sign exactly as the reference's
// double-precision lane kernel; seeded from the FP32 result (bound |d| + slack).
hipError_t launch_sdf64(const MeshDev& m, int K, int banks, const QueryArgs& a, hipStream_t s);

// CSG union scenes (scene.hip). KM = branching of the instanced meshes (all equal); banks =
// stack capacity / 64 for both the scene tree and the mesh trees. Work counters: [0] interior
// nodes, [1] triangles, [4]/[5] their wave-level fetches, [6] primitives, [7] wave-level.
hipError_t launch_scene32(const SceneDev& sc, int KM, int banks, const QueryArgs& a, bool count, hipStream_t s);
// FP64 pass over a.idx / a.count_ptr, seeded with |v32| (1 + slack_rel) + slack_abs.
hipError_t launch_scene64(const SceneDev& sc, int KM, int banks, const QueryArgs& a, hipStream_t s);

// Many meshes in one launch (batch.hip).
struct alignas(16) BatchItemDev
{
  const float* pts;
  float*       out;
  uint32_t*    tri;
  uint64_t     base;       // flat index of the item's first query
  uint32_t     n, mesh;
  float        box_lo[3];  // key box of the item's mesh: bin = (p - lo) * inv, 8 bits per axis
  float        box_inv[3];
  float        recheck_abs; // |d| below this is redone in FP64 (recheck_eps * mesh diagonal)
  float        slack_abs;   // FP64 seed slack (1e-5 * mesh diagonal)
  uint32_t     pad[2];
};
static_assert(sizeof(BatchItemDev) == 80, "BatchItemDev layout");

// A mesh's packet cost grid as the batch marking reads it (threshold 0xFFFFFFFF: no grid).
struct GridDev
{
  const uint32_t* cost; // g^3 cells, x fastest
  float           lo[3], scale[3];
  uint32_t        g, threshold, pad[2];
};
static_assert(sizeof(GridDev) == 48, "GridDev: whole 16-byte units (scalar loads)");

struct BatchArgs
{
  const BatchItemDev* items;
  const uint32_t*     packets;     // 4 per packet: item, first query (in the item's sorted order), 0, 0
  uint32_t            num_packets;
  const MeshDev*      meshes;
  const uint32_t*     sorted;      // flat query indices in sort order (null: no presort)
  uint64_t*           recheck_mask; // per packet: lane mask of its queries to redo in FP64 (null = off)
  uint32_t*           recheck_list; // packet << 6 | lane, in packet order (launch_recheck_compact)
  uint32_t*           recheck_count;
  unsigned long long* work;        // [0] inner, [1] triangles, [4]/[5] per wavefront (null = off)
  unsigned long long* pair_work;   // counting launches: [pairs, pair batches, culling loads, slab loads]
  // Early packets (early_blocks = 0: none): batch_mark lists the packets whose most expensive
  // query (the cost-grid value of its cell in its mesh's grid) reaches the mesh's threshold;
  // the first early_blocks workgroups of batch32 take them, the rest skip flagged packets.
  const GridDev*      grids;       // per mesh
  uint8_t*            flag;        // per packet
  uint32_t*           early_list;
  uint32_t*           early_count;
  uint32_t            early_cap, early_blocks;
  // Persistent launch (persist_blocks > 0): that many workgroups, each wavefront taking the next
  // entry of one queue -- the early list slots (early_cap of them when early_blocks is set),
  // then the packet table -- from *queue (zero at launch) until all queue_total are taken.
  uint32_t*           queue;
  uint32_t            queue_total, persist_blocks;
};

hipError_t launch_batch_mark(const BatchArgs& a, hipStream_t s);

// keys: uint32 when narrow (at most 256 items: 8 item bits + 24 curve bits), else uint64
hipError_t launch_batch_keys(const BatchItemDev* items, uint32_t num_items, uint64_t total, void* keys, bool narrow,
                             hipStream_t s);
// wg = 64, 128 or 256 lanes per workgroup (1, 2 or 4 packets).
hipError_t launch_batch32(const BatchArgs& a, int K, int banks, int wg, bool count, hipStream_t s);
// Workgroups of wg lanes of the persistent batch32 that fit the device at once (occupancy x CUs).
hipError_t batch32_persistent_blocks(int K, int banks, int wg, bool count, uint32_t* blocks);
// FP64 pass over the compacted recheck list (a.recheck_list == null skips); a fixed grid
// strides over groups of 64 entries.
hipError_t launch_batch64(const BatchArgs& a, int K, int banks, hipStream_t s);

// csg.hip: CSG programs. FP32 pass over a.n queries (a.idx: presort order), appending
// |v| < a.recheck_abs to the recheck list; FP64 pass over that list (a.idx, *a.count_ptr).
hipError_t launch_csg32(const CsgDev& g, int KM, int banks, const QueryArgs& a, hipStream_t s);
hipError_t launch_csg64(const CsgDev& g, int KM, int banks, const QueryArgs& a, hipStream_t s);

// cloud.hip: k nearest cloud points (k <= kKnnMax) of n query points (pts, in the order idx) or,
// with self set, of the cloud's own n points in leaf order excluding each point itself. Row r of
// out_idx / out_d2 (k entries, nearest first, missing entries kInvalid / +inf) belongs to query r
// (self: to cloud index r).
struct KnnArgs
{
  const float*    pts;
  const uint32_t* idx;
  uint64_t        n;
  uint32_t        k;
  uint32_t        self;
  uint32_t*       out_idx;
  float*          out_d2;
};
hipError_t launch_knn(const CloudDev& c, int K, int banks, const KnnArgs& a, hipStream_t s);

// util.hip
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
// v[i] <- map[v[i]] where v[i] < nmap (polygon meshes: triangle index -> face index)
hipError_t launch_map_u32(uint32_t* v, uint64_t n, const uint32_t* map, uint32_t nmap, hipStream_t s);
// Per leaf, at its first triangle: the slab half-thickness about that triangle's plane
hipError_t launch_leaf_slabs(const void* nodes, uint32_t num_nodes, int K, const float* tsph, const float* tripos,
                             float* lthick, hipStream_t s);
// Per triangle (leaf order): its record in a frame of its own plane, from the FP32 vertices
hipError_t launch_frames(const float* tripos, uint64_t nt, FrameRec* out, hipStream_t s);
// n doubles -> floats (round to nearest even) and n floats -> doubles
hipError_t launch_f64_to_f32(const double* in, float* out, uint64_t n, hipStream_t s);

// Lattice of dims[0] x dims[1] x dims[2] points origin + (i, j, k) * spacing, x fastest.
struct GridSpec
{
  double   origin[3], spacing[3];
  uint64_t dims[3];
};
hipError_t launch_grid_points(const GridSpec& g, float* out, hipStream_t s);
hipError_t launch_f32_to_f64(const float* in, double* out, uint64_t n, hipStream_t s);
hipError_t launch_add_counts(unsigned long long* work, const uint32_t* recheck, const uint32_t* fallback, hipStream_t s);

// Morton / sort (morton.hip)
// Bounding box of n points into d_box6 (lo xyz, hi xyz); part: kBboxParts*6 values of scratch.
constexpr int kBboxParts = 1024;
hipError_t launch_bbox_f32(const float* pts, uint64_t n, float* d_box6, void* part, hipStream_t s);
hipError_t launch_bbox_f64(const double* pts, uint64_t n, double* d_box6, void* part, hipStream_t s);
// codes[i] = 63-bit Morton code >> shift (0: the reference's codes; the query presort keeps the
// top 32 bits, which order 64-query packets as well and sort in half the passes)
hipError_t launch_codes_f32(const float* pts, uint64_t n, const float* d_box6, uint64_t* codes, hipStream_t s,
                            int shift = 0);
// keys[i] = 30-bit Hilbert index of the point on a 1024^3 grid of d_box6 (query presort), with
// the cost flags in bits 30 and 31 when cf is given
hipError_t launch_hilbert_f32(const float* pts, uint64_t n, const float* d_box6, uint32_t* keys, hipStream_t s,
                              const CostFlags* cf = nullptr);
hipError_t launch_codes_f64(const double* pts, uint64_t n, const double* d_box6, uint64_t* codes, hipStream_t s);
// Bytes of scratch radix_sort needs for n keys.
std::size_t radix_sort_scratch_bytes(uint64_t n);
// Stable LSD sort of (keys, values); values_in may be null (identity). Results land in
// keys_out / values_out. The 'bits' low bits of the keys are the sort key (<= 64); higher bits
// ride along unsorted.
hipError_t radix_sort_pairs(const uint64_t* keys_in, const uint32_t* values_in, uint64_t* keys_out,
                            uint32_t* values_out, uint64_t n, int bits, void* scratch, hipStream_t s);
// The same for 32-bit keys (bits <= 32); with want_keys false the sorted keys are not written
// (keys_out still serves as a ping-pong buffer).
hipError_t radix_sort_pairs32(const uint32_t* keys_in, const uint32_t* values_in, uint32_t* keys_out,
                              uint32_t* values_out, uint64_t n, int bits, void* scratch, bool want_keys,
                              hipStream_t s);

} // namespace ebgpu
