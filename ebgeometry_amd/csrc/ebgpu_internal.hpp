// Internal layouts shared by the host builder, the C-ABI layer and the HIP kernels.
//
// HBM layout of one mesh (all arrays 64-byte aligned, in leaf order where per-triangle):
//   nodes   Node<K>[num_nodes]   K child boxes as SoA (lo.x[K] .. hi.z[K]) + child refs.
//                                K=4: 128 B, K=8: 256 B. Read with wave-uniform (scalar)
//                                loads: one fetch serves all 64 queries of a wavefront.
//   tris    TriRec[num_tris]     128 B FP32 record per triangle with the quantities the
//                                point-triangle kernel needs precomputed (edges, face normal,
//                                edge x normal, 1/|edge|^2). Also read with scalar loads.
//   tripos  float[12*num_tris]   raw vertex positions (v0,v1,v2,pad) for the FP64 path.
//   tsph    float8[num_tris+16]  per-triangle culling data, contiguous per leaf: bounding sphere
//                                (centre, radius) and plane (unit normal, -n.v0), so that the
//                                culling data of 4 leaf triangles arrives in one scalar round trip.
//   feat    uint32[8*num_tris]   vertex ids, half-edge ids, original face id: lets the sign
//                                step fetch the one pseudo-normal a query needs.
//   vn32/vn64, en32/en64         vertex / half-edge pseudo-normals (float4 / double4).
//   fn64    double4[num_tris]    FP64 face normals in leaf order (FP64 recheck path).
#pragma once

#include <cstddef>
#include <cstdint>

namespace ebgpu {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr int      kMaxLeaf = 16;
constexpr int      kSphPad  = 16; // zero sphere entries after the last triangle (chunked leaf loads)

// Feature codes of the closest point on a triangle.
enum Feature : uint32_t
{
  kFace = 0,
  kVert0 = 1,
  kVert1 = 2,
  kVert2 = 3,
  kEdge0 = 4, // v0 -> v1
  kEdge1 = 5, // v1 -> v2
  kEdge2 = 6  // v2 -> v0
};

template <int K>
struct alignas(64) Node
{
  float    lox[K], loy[K], loz[K], hix[K], hiy[K], hiz[K];
  uint32_t ref[K]; // cnt==0: child node index; cnt in 1..16: first triangle of the leaf
  uint32_t cnt[K]; // 0: interior child; 1..16: leaf triangle count; kInvalid: empty slot
};
static_assert(sizeof(Node<4>) == 128, "Node<4> layout");
static_assert(sizeof(Node<8>) == 256, "Node<8> layout");

// A subtree root (node reference and box, as on the traversal stack): the roots of a frontier
// that covers the whole tree let several wavefronts share one expensive packet.
struct SplitRoot
{
  uint32_t ref, cnt; // cnt == 0: interior node ref; 1..16: leaf of cnt triangles from ref
  float    lo[3], hi[3];
};
static_assert(sizeof(SplitRoot) == 32, "SplitRoot layout");

struct alignas(64) TriRec
{
  float    v0[3];  // first vertex
  float    e0[3];  // v1 - v0   (the reference's v21)
  float    e1[3];  // v2 - v1   (v32)
  float    e2[3];  // v0 - v2   (v13)
  float    n[3];   // unit face normal (x2-x0)x(x2-x1)/|.|
  float    c0[3];  // +-e0 x n, pointing out of the triangle: sign of c0.(p-v0) tests the face region
  float    c1[3];  // e1 x n
  float    c2[3];  // e2 x n
  float    inv[3]; // 1/|e_i|^2
  uint32_t face;   // original face index
  float    sc[3];  // bounding sphere centre (float) ...
  float    sr;     // ... and radius, rounded outward: |p - sc| - sr <= distance(p, triangle)
};
static_assert(sizeof(TriRec) == 128, "TriRec layout");

// The same triangle in a frame of its own plane, for distance-only evaluation: vertices
// relabelled so that A -> B is the longest edge, ex along it, ey in the plane towards C, ez their
// cross product; in (ex, ey) A = (0, 0), B = (a, 0), C = (b, c) with c >= 0. Built in FP64.
struct alignas(64) FrameRec
{
  float o[3];      // A
  float ex[3];     // unit vectors of the frame
  float ey[3];
  float ez[3];
  float a, b, c;   // B = (a, 0), C = (b, c)
  float bma;       // b - a
  float ca;        // c * a
  float k1u, k1v;  // B -> C: t = k1u u + k1v v + k1c, ((b - a), c, -a (b - a)) / |C - B|^2 (0 for a
  float k1c;       //          zero-length edge)
  float k2u, k2v;  // A -> C: t = k2u u + k2v v, (b, c) / |C - A|^2 (0 for a zero-length edge)
  float pad[10];
};
static_assert(sizeof(FrameRec) == 128, "FrameRec layout");

// Device view of a mesh, passed by value to kernels.
struct MeshDev
{
  const void*     nodes;
  const TriRec*   tris;
  const float*    tripos; // 12 per triangle
  const float*    tsph;   // 8 floats per triangle: sphere centre, radius, normal, -n.v0 (leaf order)
  const uint32_t* feat;   // 8 per triangle: vid0..2, he0..2, face, 0
  const float*    vn32;   // float4 per vertex
  const float*    en32;   // float4 per half-edge (3*face + i)
  const double*   vn64;   // double4 per vertex
  const double*   en64;   // double4 per half-edge
  const double*   fn64;   // double4 per triangle (leaf order): FP64 face normal
  float           root_lo[3], root_hi[3];
  float           diag;
  uint32_t        num_tris;
  uint32_t*       status; // device word: bit 0 set when a traversal found its stack too small
  float           hmax;   // longest triangle edge (rounded up): scale of the FP32 distance error
  uint32_t        pad0;
  const float*    lthick; // per leaf, at its first triangle: slab half-thickness about that
                          // triangle's plane holding the whole leaf (rounded up); may be null
  const FrameRec*  frames; // per triangle (leaf order); may be null
};
static_assert(sizeof(MeshDev) % 16 == 0, "MeshDev is read with 16-byte scalar loads");
constexpr uint32_t kStatusStackOverflow = 1u;

// ---------------------------------------------------------------------------------------------
// CSG union scenes: primitives in BVH leaf order
// ---------------------------------------------------------------------------------------------
enum PrimType : uint32_t
{
  kPrimSphere = 0,
  kPrimBox    = 1,
  kPrimMesh   = 2
};

struct alignas(64) PrimRec
{
  uint32_t type, mesh, pad0, pad1;
  float    a[6];  // sphere: centre xyz, radius; box: lo xyz, hi xyz
  float    A[9];  // mesh instance: local point = A (p - t) with A = M / s, value = s * mesh(local)
  float    t[3];
  float    s;
  float    bv[6]; // world bounding box (lo, hi)
  float    pad2[3];
};
static_assert(sizeof(PrimRec) == 128, "PrimRec layout");

struct alignas(64) PrimRec64 // FP64 copy of the same parameters (near-surface recheck)
{
  double a[6];
  double A[9]; // M itself: local point = M (p - t) / s
  double t[3];
  double s;
  double pad[5];
};
static_assert(sizeof(PrimRec64) == 192, "PrimRec64 layout");

struct SceneDev
{
  const void*      nodes;  // Node<4>
  const PrimRec*   prims;
  const PrimRec64* prims64;
  const MeshDev*   meshes; // device array, indexed by PrimRec::mesh
  float            root_lo[3], root_hi[3];
  float            diag;
  uint32_t*        status; // device word: kStatusStackOverflow when the scene walk's stack was too small
};

// ---------------------------------------------------------------------------------------------
// CSG programs: an implicit-function expression evaluated per query by a small interpreter
// ---------------------------------------------------------------------------------------------
// Post-order instructions over a value stack and a point stack (the current point is the top).
// Parameters p[] in the FP32 record (CsgInstr) and the FP64 record (CsgInstr64) are the same
// numbers, rounded once.
enum CsgOpcode : uint32_t
{
  // leaves: push f(current point)
  kCsgSphere = 0,      // p0-2 centre, p3 radius
  kCsgBox,             // p0-2 lo, p3-5 hi
  kCsgPlane,           // p0-2 point, p3-5 unit normal
  kCsgTorus,           // p0-2 centre, p3 major radius, p4 minor radius
  kCsgCylinder,        // p0-2 centre, p3-5 unit axis, p6 length, p7 radius
  kCsgInfCylinder,     // p0-2 centre, p3 radius; n = axis
  kCsgCapsule,         // p0-2 first centre, p3-5 second centre, p6 radius
  kCsgInfCone,         // p0-2 tip, p3 sin(half angle), p4 cos(half angle)
  kCsgCone,            // p0-2 tip, p3 sin(half angle), p4 cos(half angle), p5 height
  kCsgRoundedBox,      // p0-2 half extents, p3 rounding radius
  kCsgRoundedCylinder, // p0 major radius, p1 minor radius, p2 half height
  kCsgMesh,            // n = mesh index: the mesh's signed distance
  // point transforms: save the current point, replace it by T(point)
  kCsgTranslate,       // q = x - p0-2
  kCsgRotate,          // n = axis, p0 cos, p1 sin (the point map of RotateIF::value)
  kCsgScalePoint,      // q = x / p0
  kCsgElongate,        // q = x - clamp(x, -e, e), e = p0-2
  kCsgReflect,         // q = x * p0-2 (each +1 or -1)
  kCsgRepeat,          // q_i = x_i - P_i round(clamp(x_i / P_i, -lo_i, hi_i)), P = p0-2, lo = p3-5, hi = p6-8
  kCsgPopPoint,        // restore the saved point
  // value ops on the top value
  kCsgScaleValue,      // v * p0
  kCsgNegate,          // -v
  kCsgOffset,          // v - p0
  kCsgAnnular,         // |v| - p0
  // accumulators: pop v and fold it into the value(s) below
  kCsgMinAcc,          // u = min(u, v)
  kCsgMaxAcc,          // u = max(u, v)
  kCsgAddAcc,          // u = u + v
  kCsgTwoMinAcc,       // (a, b) = the two smallest of a, b, v (strict <, in order)
  kCsgTwoMaxAcc,       // (a, b) = the two largest
  kCsgSmoothMinEnd,    // (a, b) -> smooth min of a and b with length p0; n = 0 quadratic, 1 exponential
  kCsgSmoothMaxEnd,    // (a, b) -> quadratic smooth max with length p0
  kCsgDifference,      // (a, b) -> max(a, -b)
  kCsgPushConst,       // push p0 (the +-inf starting values of the accumulators)
  // unions accelerated by a BVH over primitive subprograms (a = BVH index): push the value
  kCsgBvhUnion,
  kCsgBvhSmoothUnion,  // p0 smoothing length; n = 0 quadratic, 1 exponential
  kCsgNumOps
};

struct alignas(64) CsgInstr
{
  uint32_t op, n, a, b;
  float    p[12];
};
static_assert(sizeof(CsgInstr) == 64, "CsgInstr layout");

struct alignas(64) CsgInstr64
{
  double   p[12];
  uint32_t op, n, a, b;
  uint32_t pad[4];
};
static_assert(sizeof(CsgInstr64) == 128, "CsgInstr64 layout");

// One primitive of a BVH union: the subprogram ops[start, start + count) (no BVH op inside),
// leaving one value, and its bounding volume.
struct CsgPrim
{
  uint32_t start, count;
  float    bv[6];
};
static_assert(sizeof(CsgPrim) == 32, "CsgPrim layout");

struct CsgBvhDev
{
  const void*    nodes; // Node<4>, leaves index prims
  const CsgPrim* prims;
  float          root_lo[3], root_hi[3];
  uint32_t       pad[2];
};
static_assert(sizeof(CsgBvhDev) == 48, "CsgBvhDev layout");

// ---------------------------------------------------------------------------------------------
// Point clouds: K-wide tree over the points, leaves of at most 16 points
// ---------------------------------------------------------------------------------------------
struct CloudDev
{
  const void*     nodes;  // Node<K>; leaf ref = first point (leaf order), cnt = point count
  const uint32_t* pts;    // 4 words per point in leaf order: x, y, z (float bits), cloud index;
                          // padded by 16 points (chunked scalar loads)
  float           root_lo[3], root_hi[3];
  uint32_t*       status; // kStatusStackOverflow when a walk's stack was too small
};
constexpr uint32_t kKnnMax = 16; // largest k of one query

constexpr int kCsgMaxValues = 16; // value stack depth
constexpr int kCsgMaxPoints = 8;  // saved points

struct CsgDev
{
  const CsgInstr*   ops;
  const CsgInstr64* ops64;
  const CsgBvhDev*  bvhs;
  const MeshDev*    meshes;
  uint32_t          main_count; // the program is ops[0, main_count)
  uint32_t          pad;
  uint32_t*         status;     // kStatusStackOverflow: a BVH walk's stack was too small
};

} // namespace ebgpu
