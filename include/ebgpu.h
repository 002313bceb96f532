/*
 * ebgpu -- MI355X (gfx950) batched signed-distance queries: C ABI.
 *
 * Drop-in boundary for the reference's signed-distance query path. The reference exposes
 * only a per-point virtual call,
 *     ImplicitFunction<T>::value(const Vec3T<T>&) const      (Source/EBGeometry_ImplicitFunction.hpp:51-52)
 *     SignedDistanceFunction<T>::signedDistance(...)          (Source/EBGeometry_SignedDistanceFunction.hpp:54-62)
 * on objects built by
 *     TriMeshSDF(mesh, BVH::Build, maxLeafGroups)             (Source/EBGeometry_MeshDistanceFunctions.hpp:376-387)
 *     BVHUnionIF(prims, bvs)                                  (Source/EBGeometry_CSG.hpp:401-440)
 *     Translate / Rotate / Scale                              (Source/EBGeometry_Transform.hpp:44-70)
 * This header gives the batched equivalent: opaque handles built once, then queries over
 * arrays of points. Plain pointers and sizes only.
 *
 * Conventions
 *   - Signed distance is negative inside (outward-oriented triangles: the face normal is
 *     (x2-x0)x(x2-x1), Source/EBGeometry_TriangleImplem.hpp:94-106).
 *   - Points and vertices are float32 xyz triples. Distances are returned as float32.
 *   - Functions named *_host take host pointers and are synchronous. All others take device
 *     pointers and enqueue on `stream` (a hipStream_t; NULL = the null stream).
 *   - Every call returns an ebg_status; ebg_last_error() gives a message for the calling
 *     thread. Unlike the reference (which aborts on a failed EBGEOMETRY_EXPECT,
 *     Source/EBGeometry_Macros.hpp:48-63), nothing aborts.
 *   - Handles are reentrant for queries enqueued on one stream. Queries on the same handle from
 *     different streams must be ordered by the caller (they share the handle's workspace).
 */
#ifndef EBGPU_H
#define EBGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ebg_status
{
  EBG_OK             = 0,
  EBG_INVALID_ARG    = 1,
  EBG_OOM            = 2,
  EBG_HIP_ERROR      = 3,
  EBG_STACK_OVERFLOW = 4, /* a traversal needed more stack than the build guaranteed */
  EBG_NOT_READY      = 5, /* e.g. learned bound requested before a network was trained */
  EBG_UNSUPPORTED    = 6
} ebg_status;

typedef struct ebg_mesh  ebg_mesh;
typedef struct ebg_scene ebg_scene;
typedef struct ebg_batch ebg_batch;

/* Message describing the last failure on the calling thread ("" if none). */
const char* ebg_last_error(void);

/* ---------------------------------------------------------------------------------------- */
/* Mesh build -- replaces TriMeshSDF(mesh, BVH::Build::SAH|Morton, maxLeafGroups)             */
/* (Source/EBGeometry_MeshDistanceFunctionsImplem.hpp:417-460) and the DCEL pseudo-normals   */
/* it copies per triangle (Source/EBGeometry_DCEL_{Face,Edge,Vertex}Implem.hpp).              */
/* ---------------------------------------------------------------------------------------- */

enum ebg_builder
{
  EBG_BUILD_SAH    = 0, /* binned SAH on the host (default; best traversal quality)            */
  EBG_BUILD_MORTON = 1, /* Morton codes + radix sort + balanced K-ary tree on the GPU
                           (the reference's Build::Morton, Source/EBGeometry_BVHImplem.hpp:207-312);
                           leaf_size 0 or 1 gives the reference's depth floor(log n / log K),
                           larger values fewer, fuller leaves (at most 16 triangles each)     */
  EBG_BUILD_SAH_GPU = 2 /* the SAH build of EBG_BUILD_SAH (Build::SAH, Source/EBGeometry_BVH.hpp:
                           521-641) on the GPU -- the same splits, hence the same tree -- with
                           the pseudo-normals (MeshT::reconcile, DCEL_MeshImplem.hpp:304-309) and
                           the packed triangle records also computed on the device            */
};

typedef struct ebg_build_opts
{
  int32_t builder;    /* enum ebg_builder                                                   */
  int32_t branching;  /* K: 4 or 8 children per node (0 = 4)                               */
  int32_t leaf_size;  /* SAH: max triangles per leaf, 1..16 (0 = 4)                          */
  int32_t reserved;
} ebg_build_opts;

typedef struct ebg_mesh_info
{
  double   bbox_lo[3], bbox_hi[3];
  double   diag;              /* bbox diagonal                                               */
  uint64_t num_vertices, num_triangles, num_nodes, num_leaves;
  int32_t  branching, leaf_size, stack_need, stack_capacity;
  uint64_t device_bytes;      /* HBM held by the handle (excluding query workspace)          */
  double   build_seconds;
} ebg_mesh_info;

/* Triangles are (v0,v1,v2) vertex-index triples forming a closed, consistently outward
 * oriented 2-manifold (like the reference's DCEL meshes). Host pointers; data is copied. */
ebg_status ebg_mesh_create(const float* xyz, uint64_t num_vertices, const uint32_t* tri,
                           uint64_t num_triangles, const ebg_build_opts* opts, ebg_mesh** out);
/* Polygon meshes: MeshSDF<T, Meta, K>(DCEL mesh, BVH::Build) over general polygon faces
 * (Source/EBGeometry_MeshDistanceFunctions.hpp:149-281, built from a polygon soup by
 * Soup::soupToDCEL, Source/EBGeometry_SoupImplem.hpp:148-226). Face f is the vertex list
 * face_vertices[face_offsets[f] .. face_offsets[f+1]) (face_offsets has num_faces+1 entries,
 * starting at 0); faces must be planar (corners within 1e-5 of the face's extent of its plane) and
 * simple (not self-intersecting), else EBG_INVALID_ARG. The handle answers every mesh query: the
 * signed distance is FaceT::signedDistance's minimum by magnitude (DCEL_FaceImplem.hpp:388-423),
 * and closest-face outputs (d_tri) hold face indices. num_triangles in ebg_mesh_info counts the
 * triangles the faces are clipped into. */
ebg_status ebg_mesh_create_polygons(const float* xyz, uint64_t num_vertices, const uint32_t* face_offsets,
                                    const uint32_t* face_vertices, uint64_t num_faces, const ebg_build_opts* opts,
                                    ebg_mesh** out);
/* The face each of the handle's num_triangles triangles belongs to (identity for triangle meshes). */
ebg_status ebg_mesh_triangle_faces(const ebg_mesh* mesh, uint32_t* tri_face);
ebg_status ebg_mesh_destroy(ebg_mesh* mesh);
ebg_status ebg_mesh_get_info(const ebg_mesh* mesh, ebg_mesh_info* info);
/* Original face index of each triangle in BVH leaf order (num_triangles uint32, host). For the
 * Morton builder this is the stable (code, index) sort of the triangle box centroids' Morton
 * codes, the order SFC::order gives up to equal codes (Source/EBGeometry_SFCImplem.hpp:279-300). */
ebg_status ebg_mesh_leaf_order(const ebg_mesh* mesh, uint32_t* faces);

/* Host-only (no GPU needed): the pseudo-normals ebg_mesh_create computes, per face: face normal
 * (3 doubles), vertex normals of v0,v1,v2 (9), normals of the edges v0->v1, v1->v2, v2->v0 (9).
 * Replaces MeshT::reconcile(VertexNormalWeight::Angle) (Source/EBGeometry_DCEL_MeshImplem.hpp:304-309)
 * as it is read back per triangle by TriMeshSDF (MeshDistanceFunctionsImplem.hpp:429-452). */
ebg_status ebg_pseudo_normals_host(const float* xyz, uint64_t num_vertices, const uint32_t* tri,
                                   uint64_t num_triangles, double* face_n, double* vert_n,
                                   double* edge_n);

/* Host-only (no GPU needed): the triangles ebg_mesh_create_polygons clips the faces into (each
 * face of n corners gives n-2; tri: 3 per triangle, tri_face: its face) and their pseudo-normals in
 * the layout of ebg_pseudo_normals_host: the face's normal, the polygon mesh's vertex normals, and
 * per triangle edge the face half-edge's normal, or the face normal for an edge inside the face.
 * Counterpart of the normals MeshT::reconcile gives a polygon mesh
 * (Source/EBGeometry_DCEL_MeshImplem.hpp:304-309). */
ebg_status ebg_polygons_to_triangles_host(const float* xyz, uint64_t num_vertices, const uint32_t* face_offsets,
                                         const uint32_t* face_vertices, uint64_t num_faces, uint32_t* tri,
                                         uint32_t* tri_face, double* face_n, double* vert_n, double* edge_n);

/* ---------------------------------------------------------------------------------------- */
/* Queries -- batched TriMeshSDF::signedDistance / getClosestTriangle                         */
/* (Source/EBGeometry_MeshDistanceFunctionsImplem.hpp:480-549)                                */
/* ---------------------------------------------------------------------------------------- */

typedef struct ebg_query_opts
{
  int32_t  presort;        /* 1: sort queries by Morton code before traversal (default)     */
  int32_t  fp64_recheck;   /* 1: redo |d| < recheck_eps*diag in FP64 (default)              */
  double   recheck_eps;    /* default 1e-4                                                  */
  int32_t  learned_bound;  /* 1: seed each query's cull radius with the mesh's trained MLP */
  int32_t  fp64_all;       /* 1: FP64 polish of every query (exact mode, slower)            */
  float    bound_scale;    /* radius = bound_scale*pred + bound_offset*diag (1.25, 1e-3)    */
  float    bound_offset;
  void*    stream;         /* hipStream_t                                                   */
  uint64_t* work;          /* optional device array of 8 uint64 accumulated per call:
                              [0] interior nodes whose children each query tested,
                              [1] real triangles tested by each query,
                              [2] queries rechecked in FP64,
                              [3] queries that fell back to the full traversal,
                              [4] interior node records fetched per wavefront (all waves),
                              [5] triangle records fetched per wavefront (all launches),
                              [6], [7] as [0], [1] for the unseeded re-run of fallbacks
                              ([0], [1] then cover the seeded launch only).
                              Counting uses a slower kernel variant and splits no packets
                              (the counts are the plain traversal's); time without it.      */
  const float* seed_radius; /* optional device array (indexed by query): a caller-supplied
                              cull radius per query, used like the learned bound (queries
                              with no triangle inside it are redone unseeded). NULL = none. */
  uint64_t* pair_work;     /* optional device array of 4 uint64, accumulated by counting calls
                              (work != NULL) of the FP32 distance-only mesh queries: [0] the
                              (query, triangle) pairs evaluated by the pair-compacted leaves,
                              [1] their 64-lane evaluation batches, [2] the screen's loads of
                              culling records (4 triangles, 128 B each, per wavefront), [3] the
                              leaves whose slab was tested (one 4-B load each, per wavefront).
                              NULL = not counted.                                            */
} ebg_query_opts;

/* Fill *opts with the defaults above. */
void ebg_query_opts_default(ebg_query_opts* opts);

/* d_pts: n float32 xyz triples; d_dist: n float32; d_tri (nullable): n uint32 face indices
 * (closest face, up to exact ties). Device pointers, asynchronous on opts->stream.
 * Split packets: with presort on and n >= 2^14, packets of 64 queries whose estimated cost (a
 * 32^3 grid of measured packet costs over the mesh box, built at the first such call) is a few
 * times the grid mean are each shared among 16 wavefronts walking disjoint subtrees, unseeded,
 * inside the same traversal launch; results are those of the plain traversal. */
ebg_status ebg_query_sdf(const ebg_mesh* mesh, const float* d_pts, uint64_t n, float* d_dist,
                         uint32_t* d_tri, const ebg_query_opts* opts);
/* Allocation-free form of ebg_query_sdf. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). Same arguments and results as ebg_query_sdf; it calls no hipMalloc, hipFree
 * or other allocating routine and creates no stream or event. Its workspace is one fixed device
 * array of 40 MiB per GPU, held from the library's first use on that GPU until the process ends,
 * which limits a call to n <= EBG_NOALLOC_MAX_QUERIES (larger n: EBG_INVALID_ARG). Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller. Differences from ebg_query_sdf: split packets are used
 * only once an ebg_query_sdf call of >= 2^14 presorted queries has built the mesh's cost grid;
 * the learned bound's network runs on opts->stream before the traversal (no second stream); the
 * traversal is not timed (ebg_mesh_last_traversal_ms keeps the previous value). */
#define EBG_NOALLOC_MAX_QUERIES 1048576u
ebg_status ebg_query_sdf_noalloc(const ebg_mesh* mesh, const float* d_pts, uint64_t n, float* d_dist,
                                 uint32_t* d_tri, const ebg_query_opts* opts);
/* FP64 points and distances (the reference's T = double interface, SignedDistanceFunction<double>,
 * Source/EBGeometry_SignedDistanceFunction.hpp:54-62): the points are rounded to FP32 on the
 * device (round to nearest even), queried as by ebg_query_sdf, and the distances widened to
 * double. Results are those of the rounded points: within 1e-5 diag of the reference's FP64
 * value at the rounded point. Device pointers, asynchronous on opts->stream; uses a buffer of
 * 16 bytes per query held by the mesh. */
ebg_status ebg_query_sdf_f64(const ebg_mesh* mesh, const double* d_pts, uint64_t n, double* d_dist,
                             uint32_t* d_tri, const ebg_query_opts* opts);
/* Host-pointer convenience form; synchronous. */
ebg_status ebg_query_sdf_host(const ebg_mesh* mesh, const float* pts, uint64_t n, float* dist,
                              uint32_t* tri, const ebg_query_opts* opts);

/* Lattice queries for embedded-boundary generators: the signed distance at every point
 * origin + (i, j, k) * spacing, 0 <= i < dims[0], 0 <= j < dims[1], 0 <= k < dims[2], written
 * x fastest (d_dist[(k * dims[1] + j) * dims[0] + i]: the layout of an AMReX or Chombo FAB over
 * the box). Replaces the per-cell / per-node value() calls an EB generator makes through the
 * reference (AMReXSDF::operator(), Integrations/AMReX/MeshSDF/main.cpp:23-69, called by
 * EB2::GeometryShop for every node; ChomboSDF::value, Integrations/Chombo/MeshSDF/main.cpp:21-50):
 * one call per box. The points are made on the device (each coordinate the FP64 product
 * i * spacing plus origin, each rounded, then rounded to FP32), then queried as by
 * ebg_query_sdf (presorted, FP64 recheck near the surface). At most 2^32-1 points per call;
 * uses a buffer of 12 bytes per point held by the mesh. Device output, asynchronous on
 * opts->stream; the _host form copies to a host array and synchronises. */
ebg_status ebg_query_grid(const ebg_mesh* mesh, const double origin[3], const double spacing[3],
                          const uint64_t dims[3], float* d_dist, const ebg_query_opts* opts);
ebg_status ebg_query_grid_host(const ebg_mesh* mesh, const double origin[3], const double spacing[3],
                               const uint64_t dims[3], float* dist, const ebg_query_opts* opts);

/* Time the dominant traversal kernel of the last query on `mesh` (HIP events recorded on the
 * launch stream around the FP32 traversal -- the launch, plain or seeded, with its split
 * packets, and their resolve; the FP64 pass is not included). Milliseconds, or <0 if none.
 * Waits for that launch. */
double ebg_mesh_last_traversal_ms(const ebg_mesh* mesh);
/* Waits for the device, then reports whether any traversal of this mesh since its creation
 * (alone, in a batch or instanced in a scene) needed more traversal stack than the build-time
 * worst case the handle was sized for: EBG_STACK_OVERFLOW (those queries' results are not
 * exact), else EBG_OK. The capacity is the tree's exact worst case, so this is a guard, not an
 * expected outcome. No counterpart in the reference (its stack is a fixed 256 entries,
 * Source/EBGeometry_BVHImplem.hpp:1138). */
ebg_status ebg_mesh_check(const ebg_mesh* mesh);
/* Number of split packets of the last query on `mesh` (0 if none were split, <0 on error).
 * Diagnostic with no counterpart in the reference; waits for the device. */
int64_t ebg_mesh_last_split_packets(const ebg_mesh* mesh);

/* ---------------------------------------------------------------------------------------- */
/* CSG union scenes -- BVHUnionIF(primitives, bounding volumes) and its value()               */
/* (Source/EBGeometry_CSG.hpp:401-440, CSGImplem.hpp:51-95, 466-501) over SphereSDF and      */
/* BoxSDF (AnalyticDistanceFunctions.hpp:137-423) and mesh instances                          */
/* Translate(Rotate(Rotate(Rotate(Scale(mesh, s), ax, 0), ay, 1), az, 2), t)                  */
/* (Transform.hpp:44-70, TransformImplem.hpp:140-270).                                        */
/* ---------------------------------------------------------------------------------------- */

typedef enum ebg_prim_type
{
  EBG_PRIM_SPHERE = 0, /* p[0..2] centre, p[3] radius                                         */
  EBG_PRIM_BOX    = 1, /* p[0..2] low corner, p[3..5] high corner (axis-aligned)              */
  EBG_PRIM_MESH   = 2  /* meshes[mesh] under the similarity map below                          */
} ebg_prim_type;

typedef struct ebg_prim
{
  int32_t type;          /* ebg_prim_type                                                    */
  int32_t mesh;          /* EBG_PRIM_MESH: index into the scene's mesh array                 */
  double  p[6];          /* sphere / box parameters                                          */
  double  translate[3];  /* EBG_PRIM_MESH: value(x) = scale * mesh(M (x - t) / scale), the   */
  double  matrix[9];     /*   form every chain of Translate / Rotate / Scale takes. M is     */
                         /*   orthonormal, row-major (all zeros = identity); for            */
                         /*   Translate(Rz(Ry(Rx(Scale(mesh))))) it is Rx Ry Rz with R_axis */
                         /*   the point map of RotateIF::value.                             */
  double  scale;         /*   uniform scale > 0                                               */
  double  bv_lo[3];      /* bounding volume of the primitive, as passed to BVHUnionIF; it    */
  double  bv_hi[3];      /*   must contain the primitive. bv_lo[0] > bv_hi[0] = compute it:  */
                         /*   exact for spheres and boxes, the transformed mesh bbox corners */
                         /*   for mesh instances.                                            */
} ebg_prim;

/* Build a scene over n primitives. meshes[0..num_meshes) are retained by the scene (a mesh
 * handle may be destroyed by the caller at any time; its memory lives until the last scene
 * using it is destroyed). All meshes must have the same branching factor. opts->leaf_size
 * sets primitives per BVH leaf (default 4); builder must be EBG_BUILD_SAH. */
ebg_status ebg_scene_create(const ebg_prim* prims, uint64_t n, ebg_mesh* const* meshes, uint64_t num_meshes,
                            const ebg_build_opts* opts, ebg_scene** out);
ebg_status ebg_scene_destroy(ebg_scene* scene);
/* World bounding box (lo, hi) of the scene and its diagonal. */
ebg_status ebg_scene_bounds(const ebg_scene* scene, double* lo3, double* hi3, double* diag);

/* Batched BVHUnionIF::value: d_pts n float32 xyz, d_value n float32 (device pointers, async on
 * opts->stream). Honours presort, fp64_recheck / recheck_eps (relative to the scene diagonal),
 * fp64_all, stream and work ([0] interior nodes incl. mesh nodes, [1] triangles, [2] rechecked,
 * [4]/[5] wave-level fetches, [6] primitives evaluated, [7] wave-level primitive fetches; with a
 * seeded walk [3] = queries redone unseeded). learned_bound (after ebg_scene_mlp_train) or
 * seed_radius seeds each query's walk with a radius, the value-based pruning of
 * BVHUnionIF::value (Source/EBGeometry_CSGImplem.hpp:490-499) starting from it; a query none of
 * whose primitives comes in below its radius is redone unseeded, so the result is the
 * unseeded one either way. */
ebg_status ebg_scene_query(const ebg_scene* scene, const float* d_pts, uint64_t n, float* d_value,
                           const ebg_query_opts* opts);
ebg_status ebg_scene_query_host(const ebg_scene* scene, const float* pts, uint64_t n, float* value,
                                const ebg_query_opts* opts);
/* Lattice form of ebg_scene_query (same lattice and layout as ebg_query_grid). */
ebg_status ebg_scene_query_grid(const ebg_scene* scene, const double origin[3], const double spacing[3],
                                const uint64_t dims[3], float* d_value, const ebg_query_opts* opts);
ebg_status ebg_scene_query_grid_host(const ebg_scene* scene, const double origin[3],
                                     const double spacing[3], const uint64_t dims[3], float* value,
                                     const ebg_query_opts* opts);
double     ebg_scene_last_traversal_ms(const ebg_scene* scene);
/* As ebg_mesh_check, for the scene's own tree walk (its meshes report through ebg_mesh_check). */
ebg_status ebg_scene_check(const ebg_scene* scene);

/* ---------------------------------------------------------------------------------------- */
/* CSG programs -- the reference's implicit-function algebra on the GPU:                      */
/*   analytic SDFs  Plane, Sphere, Box, Torus, Cylinder, InfiniteCylinder, Capsule,            */
/*                  InfiniteCone, Cone, RoundedBox, RoundedCylinder                            */
/*                  (Source/EBGeometry_AnalyticDistanceFunctions.hpp:41-1750), and meshes;     */
/*   transforms     Translate, Rotate, Scale, Elongate, Reflect, Complement, Offset, Annular,   */
/*                  Blur, Mollify (Source/EBGeometry_Transform.hpp:78-934);                    */
/*   CSG            Union, SmoothUnion, Intersection, SmoothIntersection, Difference,          */
/*                  SmoothDifference, FiniteRepetition, BVHUnion, BVHSmoothUnion               */
/*                  (Source/EBGeometry_CSG.hpp:40-866, CSGImplem.hpp).                         */
/* An expression is given as the post-order list of its instructions (a stack program): each  */
/* leaf pushes its value at the current point; a point transform replaces the current point   */
/* until the matching EBG_CSG_POP_POINT; value transforms change the top value; folds pop the  */
/* top value into the one(s) below. The Python package (ebgeometry_amd.csg) compiles the      */
/* reference's object names into this list.                                                   */
/* ---------------------------------------------------------------------------------------- */

typedef struct ebg_csg ebg_csg;

enum ebg_csg_opcode
{
  /* leaves: push f(current point); p[] as listed */
  EBG_CSG_SPHERE = 0,      /* p0-2 centre, p3 radius                                              */
  EBG_CSG_BOX,             /* p0-2 low corner, p3-5 high corner                                   */
  EBG_CSG_PLANE,           /* p0-2 point, p3-5 unit normal                                        */
  EBG_CSG_TORUS,           /* p0-2 centre, p3 major radius, p4 minor radius                       */
  EBG_CSG_CYLINDER,        /* p0-2 centre, p3-5 unit axis, p6 length, p7 radius                   */
  EBG_CSG_INF_CYLINDER,    /* p0-2 centre, p3 radius; n = axis (0, 1, 2)                          */
  EBG_CSG_CAPSULE,         /* p0-2, p3-5 the two sphere centres, p6 radius                        */
  EBG_CSG_INF_CONE,        /* p0-2 tip, p3 sin and p4 cos of the half opening angle               */
  EBG_CSG_CONE,            /* p0-2 tip, p3 sin, p4 cos of the half angle, p5 height               */
  EBG_CSG_ROUNDED_BOX,     /* p0-2 half extents, p3 rounding radius                               */
  EBG_CSG_ROUNDED_CYLINDER,/* p0 radius - curvature, p1 curvature, p2 height / 2 - curvature      */
  EBG_CSG_MESH,            /* n = index into the program's meshes: their signed distance          */
  /* point transforms: the current point x becomes q until EBG_CSG_POP_POINT */
  EBG_CSG_TRANSLATE,       /* q = x - p0-2                                                        */
  EBG_CSG_ROTATE,          /* n = axis, p0 = cos, p1 = sin of the angle (RotateIF's point map)    */
  EBG_CSG_SCALE_POINT,     /* q = x / p0 (with EBG_CSG_SCALE_VALUE p0 after: ScaleIF)             */
  EBG_CSG_ELONGATE,        /* q = x - clamp(x, -p0-2, p0-2)                                       */
  EBG_CSG_REFLECT,         /* q = x * p0-2 (each +1 or -1)                                        */
  EBG_CSG_REPEAT,          /* q_i = x_i - P_i round(clamp(x_i / P_i, -lo_i, hi_i)), P = p0-2,
                              lo = p3-5, hi = p6-8 (FiniteRepetitionIF)                          */
  EBG_CSG_POP_POINT,
  /* value transforms of the top value */
  EBG_CSG_SCALE_VALUE,     /* v * p0                                                              */
  EBG_CSG_NEGATE,          /* -v (ComplementIF)                                                   */
  EBG_CSG_OFFSET,          /* v - p0 (OffsetIF)                                                   */
  EBG_CSG_ANNULAR,         /* |v| - p0 (AnnularIF)                                                */
  /* folds: pop v into the value(s) below */
  EBG_CSG_MIN_ACC,         /* u = min(u, v)                                                       */
  EBG_CSG_MAX_ACC,         /* u = max(u, v)                                                       */
  EBG_CSG_ADD_ACC,         /* u = u + v                                                           */
  EBG_CSG_TWO_MIN_ACC,     /* (a, b) = two smallest of a, b, v (SmoothUnionIF's scan)             */
  EBG_CSG_TWO_MAX_ACC,     /* (a, b) = two largest (SmoothIntersectionIF's scan)                  */
  EBG_CSG_SMOOTH_MIN_END,  /* (a, b) -> smooth min, length p0; n = 0 SmoothMin, 1 ExpMin          */
  EBG_CSG_SMOOTH_MAX_END,  /* (a, b) -> SmoothMax, length p0                                      */
  EBG_CSG_DIFFERENCE,      /* (a, b) -> max(a, -b)                                                */
  EBG_CSG_PUSH_CONST,      /* push p0 (may be +-inf: the starting values of the folds)            */
  /* BVH unions: a = first primitive, b = primitive count; push min over the primitives        */
  EBG_CSG_BVH_UNION,
  EBG_CSG_BVH_SMOOTH_UNION /* smooth min of the two smallest, length p0; n = 0 or 1 as above      */
};

typedef struct ebg_csg_op
{
  int32_t op, n, a, b;
  double  p[12];
} ebg_csg_op;

/* A primitive of a BVH union: the subprogram ops[start, start + count), which leaves one value
 * and contains no BVH union, and its bounding volume (the AABB the reference's BVHUnionIF takes
 * per primitive: the function's value at a point is at least the point's distance to it). */
typedef struct ebg_csg_prim
{
  uint32_t start, count;
  double   bv_lo[3], bv_hi[3];
} ebg_csg_prim;

/* The program is ops[0, main_count); ops beyond are subprograms of BVH-union primitives. All
 * meshes must share their branching factor; they are retained by the program. length_scale sets
 * the FP64 recheck threshold (recheck_eps * length_scale; 0 = 1). At most 16 stack values and 8
 * saved points at any time (EBG_UNSUPPORTED beyond). */
ebg_status ebg_csg_create(const ebg_csg_op* ops, uint64_t num_ops, uint64_t main_count,
                          const ebg_csg_prim* prims, uint64_t num_prims, ebg_mesh* const* meshes,
                          uint64_t num_meshes, double length_scale, ebg_csg** out);
ebg_status ebg_csg_destroy(ebg_csg* program);
/* Value at n points (device arrays, asynchronous on opts->stream): FP32 evaluation of every
 * query, FP64 re-evaluation of those with |v| < recheck_eps * length_scale. */
ebg_status ebg_csg_query(const ebg_csg* program, const float* d_pts, uint64_t n, float* d_value,
                         const ebg_query_opts* opts);
ebg_status ebg_csg_query_host(const ebg_csg* program, const float* pts, uint64_t n, float* value,
                              const ebg_query_opts* opts);
/* Lattice form (same lattice and layout as ebg_query_grid). */
ebg_status ebg_csg_query_grid(const ebg_csg* program, const double origin[3], const double spacing[3],
                              const uint64_t dims[3], float* d_value, const ebg_query_opts* opts);
ebg_status ebg_csg_query_grid_host(const ebg_csg* program, const double origin[3],
                                   const double spacing[3], const uint64_t dims[3], float* value,
                                   const ebg_query_opts* opts);
/* EBG_STACK_OVERFLOW if a BVH walk of this program ever exceeded its build-time stack bound. */
ebg_status ebg_csg_check(const ebg_csg* program);

/* ---------------------------------------------------------------------------------------- */
/* Point clouds -- PointCloudBVH / PointCloudHashGrid queries (Source/EBGeometry_PointCloudBVH  */
/* .hpp:115-161, PointCloudHashGrid.hpp:91-135): the k nearest cloud points of query points    */
/* (closestPoint, closestPoints) and of the cloud's own points excluding themselves           */
/* (nearestNeighbor(s), allNearestNeighbors), k = 1..16, by squared Euclidean distance in FP32. */
/* Results are sorted nearest first; equal distances keep the order in which the walk met the */
/* points (any order of exact ties is a valid answer). Missing entries (fewer than k points)   */
/* are index 0xFFFFFFFF with squared distance +inf.                                            */
/* ---------------------------------------------------------------------------------------- */

typedef struct ebg_cloud ebg_cloud;

/* n float32 xyz points (host array, copied); opts: branching 4 or 8, leaf_size 1..16 points
 * (0 = 8); the builder must be EBG_BUILD_SAH. Cloud index i is the i-th input point. */
ebg_status ebg_cloud_create(const float* xyz, uint64_t n, const ebg_build_opts* opts, ebg_cloud** out);
ebg_status ebg_cloud_destroy(ebg_cloud* cloud);
/* k nearest cloud points of n query points: d_index[q * k + j], d_dist2 (nullable) likewise.
 * Device arrays, asynchronous on opts->stream (presorted along a Hilbert curve unless
 * opts->presort is 0). */
ebg_status ebg_cloud_knn(const ebg_cloud* cloud, const float* d_pts, uint64_t n, uint32_t k,
                         uint32_t* d_index, float* d_dist2, const ebg_query_opts* opts);
ebg_status ebg_cloud_knn_host(const ebg_cloud* cloud, const float* pts, uint64_t n, uint32_t k,
                              uint32_t* index, float* dist2, const ebg_query_opts* opts);
/* k nearest other points of every cloud point (row i: cloud point i; its own index excluded,
 * coincident points with other indices included). */
ebg_status ebg_cloud_knn_self(const ebg_cloud* cloud, uint32_t k, uint32_t* d_index, float* d_dist2,
                              const ebg_query_opts* opts);
ebg_status ebg_cloud_knn_self_host(const ebg_cloud* cloud, uint32_t k, uint32_t* index, float* dist2,
                                   const ebg_query_opts* opts);
/* EBG_STACK_OVERFLOW if a walk of this cloud ever exceeded its build-time stack bound. */
ebg_status ebg_cloud_check(const ebg_cloud* cloud);

/* ---------------------------------------------------------------------------------------- */
/* Batched queries over many meshes -- many TriMeshSDF objects (MeshDistanceFunctions.hpp:    */
/* 376-387) queried by one traversal launch. The reference has no counterpart: its callers    */
/* loop over meshes and points. A batch owns its meshes (built from raw arrays, optionally    */
/* with the branching factor and leaf size picked by timing), a plan fixes one set of query   */
/* arrays and replays the whole pipeline (presort, traversal, FP64 pass) as one hipGraph.     */
/* ---------------------------------------------------------------------------------------- */

typedef struct ebg_batch_plan ebg_batch_plan;

typedef struct ebg_batch_opts
{
  int32_t autotune;   /* 1: time (branching, leaf_size) in {4,8}x{8,16} on the largest meshes
                         with synthetic queries and build every mesh with the fastest         */
  int32_t branching;  /* used when autotune = 0 (0 = 4)                                       */
  int32_t leaf_size;  /* used when autotune = 0 (0 = 8)                                       */
  int32_t threads;    /* host threads for building meshes (0 = hardware concurrency)          */
} ebg_batch_opts;

typedef struct ebg_batch_info
{
  uint64_t num_meshes, num_triangles;
  int32_t  branching, leaf_size, stack_capacity;
  int32_t  tuned;                /* 1 if the autotuner chose branching / leaf_size           */
  double   tune_ms[4];           /* traversal ms of the candidates (4,8) (4,16) (8,8) (8,16)  */
  double   build_seconds;
} ebg_batch_info;

/* One query array against one mesh of the batch (device pointers). */
typedef struct ebg_batch_item
{
  uint32_t     mesh;     /* index into the batch's meshes                                    */
  uint32_t     reserved;
  uint64_t     n;        /* < 2^32                                                            */
  const float* d_pts;    /* n float32 xyz                                                     */
  float*       d_dist;   /* n float32 signed distances                                        */
  uint32_t*    d_tri;    /* nullable: closest face per query                                  */
} ebg_batch_item;

/* meshes: num_meshes (xyz, nv, tri, nt) arrays as ebg_mesh_create takes them (host, copied). */
ebg_status ebg_batch_create(const float* const* xyz, const uint64_t* nv, const uint32_t* const* tri,
                            const uint64_t* nt, uint64_t num_meshes, const ebg_batch_opts* opts,
                            ebg_batch** out);
ebg_status ebg_batch_destroy(ebg_batch* batch);
ebg_status ebg_batch_get_info(const ebg_batch* batch, ebg_batch_info* info);
/* The batch's mesh i as a mesh handle (owned by the batch; valid while the batch lives). */
ebg_status ebg_batch_mesh(const ebg_batch* batch, uint64_t i, ebg_mesh** mesh);

/* Launch shape flag of ebg_batch_plan_create's `workgroup`: a persistent grid (as many
 * workgroups as fit the device at once) whose wavefronts take packets from a queue, in the
 * order of the one-wavefront-per-packet launch, until none is left. */
#define EBG_PLAN_PERSISTENT 0x10000

/* Plan one set of items: checks them, builds the packet schedule (64-query packets of one item
 * each, largest meshes first), allocates every buffer, and -- if `workgroup` is 0 -- picks the
 * launch shape (one wavefront per packet in workgroups of 64, 128 or 256 lanes, or the persistent
 * grid with 64 or 256) by timing the traversal on the items as they are (their points must be
 * valid then); otherwise `workgroup` is 64, 128 or 256, optionally | EBG_PLAN_PERSISTENT. With capture = 1 the pipeline is
 * recorded once into a hipGraph; ebg_batch_plan_launch then replays it. The item arrays must
 * stay allocated while the plan is used; their contents may change between launches.
 * opts: presort, fp64_recheck, recheck_eps, fp64_all, work are honoured; stream is used for
 * tuning; learned_bound and seed_radius must be off. */
ebg_status ebg_batch_plan_create(ebg_batch* batch, const ebg_batch_item* items, uint64_t num_items,
                                 const ebg_query_opts* opts, int32_t workgroup, int32_t capture,
                                 ebg_batch_plan** out);
ebg_status ebg_batch_plan_launch(ebg_batch_plan* plan, void* stream);
ebg_status ebg_batch_plan_destroy(ebg_batch_plan* plan);
/* Launch shape the plan uses (workgroup lanes | EBG_PLAN_PERSISTENT), and the time of its last launch (HIP events on the launch
 * stream; waits for it): for a plan without a graph, around the traversal kernel alone; for a
 * captured plan, around the whole graph replay (keys, sort, traversal, FP64 pass) -- events
 * recorded inside a graph are not timed. */
int32_t    ebg_batch_plan_workgroup(const ebg_batch_plan* plan);
double     ebg_batch_plan_last_traversal_ms(const ebg_batch_plan* plan);
/* Early packets of the last launch: packets of meshes with at least 16,384 faces whose
 * estimated cost (the mesh's packet cost grid, built once per mesh at plan creation) is >= 2x
 * the mesh's grid mean are started first. Diagnostic with no counterpart in the reference;
 * waits for the device; 0 if none, <0 on error. */
int64_t    ebg_batch_plan_last_early_packets(const ebg_batch_plan* plan);

/* One-shot convenience: plan (no tuning, no graph), launch on opts->stream, release. */
ebg_status ebg_batch_query(ebg_batch* batch, const ebg_batch_item* items, uint64_t num_items,
                           const ebg_query_opts* opts);

/* ---------------------------------------------------------------------------------------- */
/* Space-filling curve -- SFC::computeBins<T>, Morton::encode, SFC::order                      */
/* (Source/EBGeometry_SFCImplem.hpp:31-58, 228-300). Device pointers, async on stream.        */
/* Bit-exact with the reference; equal codes are ordered by increasing index (stable).        */
/* ---------------------------------------------------------------------------------------- */

ebg_status ebg_morton_codes_f32(const float* d_pts, uint64_t n, uint64_t* d_codes, void* stream);
ebg_status ebg_morton_codes_f64(const double* d_pts, uint64_t n, uint64_t* d_codes, void* stream);
/* Stable LSD radix sort of (code, index) pairs; d_perm receives the sorted indices. Scratch is
 * allocated internally. d_sorted_codes is nullable. */
ebg_status ebg_morton_sort(const uint64_t* d_codes, uint64_t n, uint32_t* d_perm,
                           uint64_t* d_sorted_codes, void* stream);

/* ---------------------------------------------------------------------------------------- */
/* Learned cull bound (no reference counterpart)                                              */
/* ---------------------------------------------------------------------------------------- */

typedef struct ebg_mlp_opts
{
  uint64_t num_samples;  /* training pairs generated by the mesh's own BVH (default 4M)      */
  int32_t  epochs;       /* default 8                                                       */
  int32_t  batch;        /* default 4096                                                    */
  float    lr;           /* Adam learning rate (default 2e-3)                               */
  uint64_t seed;
  float    near_sigma;   /* surface offset sigma for half the samples (default 0.02)         */
  float    box_scale;    /* uniform half sampled in box_scale x bbox (default 1.2)            */
  void*    stream;
  float    under_weight; /* loss weight of under-predictions (default 16): higher = fewer
                            fallbacks, looser radii                                          */
  int32_t  reserved;
} ebg_mlp_opts;

void       ebg_mlp_opts_default(ebg_mlp_opts* opts);
/* Train the mesh's bound network on the GPU (BF16 MFMA forward/backward, Adam). Writes the
 * final mean training loss to *loss (nullable). */
ebg_status ebg_mlp_train(ebg_mesh* mesh, const ebg_mlp_opts* opts, float* loss);
/* Forward pass only: predicted unsigned distance for n device points. */
ebg_status ebg_mlp_predict(const ebg_mesh* mesh, const float* d_pts, uint64_t n, float* d_pred,
                           void* stream);
/* The same network for a union scene (C5's per-scene bound): trained on num_samples points, half
 * uniform in box_scale x the scene box, half uniform in a random primitive's world box grown by
 * near_sigma x the scene diagonal, labelled |value| by the scene's own walk
 * (ebg_scene_query); arms ebg_query_opts.learned_bound for ebg_scene_query. */
ebg_status ebg_scene_mlp_train(ebg_scene* scene, const ebg_mlp_opts* opts, float* loss);
ebg_status ebg_scene_mlp_predict(const ebg_scene* scene, const float* d_pts, uint64_t n, float* d_pred,
                                 void* stream);
/* Parameters as float32 in a fixed order (see ebg_mlp_param_count); host pointers:
 * W1 (51x128, [in][out] row-major), b1 (128), W2, b2, W3, b3, W4, b4 (128x128 / 128 each),
 * w5 (128), b5 (1). set_params also arms the learned bound (it counts as trained). */
uint64_t   ebg_mlp_param_count(void);
ebg_status ebg_mlp_get_params(const ebg_mesh* mesh, float* params);
ebg_status ebg_mlp_set_params(ebg_mesh* mesh, const float* params);
/* Save the mesh's trained network (parameters, input box, loss settings) to a file, tagged with a
 * 64-bit hash of the vertex and triangle arrays the mesh was created from; ebg_mlp_load reads it
 * back into a mesh created from the same arrays (another mesh: EBG_INVALID_ARG) and arms the
 * learned bound, so a network is trained once per mesh. No counterpart in the reference. */
ebg_status ebg_mlp_save(const ebg_mesh* mesh, const char* path);
ebg_status ebg_mlp_load(ebg_mesh* mesh, const char* path);
/* Host-only check (no GPU needed, nothing is allocated or launched): the launch shapes and the
 * largest index of every buffer access of each training step ebg_mlp_train would enqueue for
 * num_samples pairs in batches of `batch` (full batches and the last, partial one) against the
 * buffers it allocates. EBG_INVALID_ARG names the failing check. ebg_mlp_train runs it first. */
ebg_status ebg_mlp_check_schedule(uint64_t num_samples, int32_t batch);
/* Training loss summed over n <= 65536 device points with device targets |d| (d_loss, one
 * float, nullable) and its gradient (d_grad, ebg_mlp_param_count floats, parameter order
 * above). The loss is the one the trainer minimises: the network output y estimates
 * log(|d|/diag + 0.005); e = y - log(|d|/diag + 0.005), loss = w e^2 with w = under_weight
 * (last trained value, default 8) where e < 0 and 1 elsewhere; the gradient is that of the mean
 * over the batch. For checking the backward kernels. */
ebg_status ebg_mlp_loss_grad(ebg_mesh* mesh, const float* d_pts, const float* d_target, uint64_t n,
                             float* d_grad, float* d_loss, void* stream);

#ifdef __cplusplus
}
#endif
#endif
