/*
 * CPU oracle for the batched signed-distance path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, double-precision restatement of the reference algorithm. Only tests/,
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py may load it, and only
 * as the checker; the product (ebgeometry_amd / libebgpu) never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to the
 * reference root). How each part is pinned (golden fixtures written by the compiled
 * reference, closed forms, brute force, invariants) is listed per function; nothing
 * here is "parity unpinned".
 */
#ifndef EBG_ORACLE_H
#define EBG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh  orc_mesh;
typedef struct orc_scene orc_scene;

/* Pseudo-normals and mesh SDF --------------------------------------------------------- */

/* Build the oracle's triangle mesh SDF: DCEL-equivalent pseudo-normals
 * (Source/EBGeometry_DCEL_FaceImplem.hpp:138-165, DCEL_EdgeImplem.hpp:130-165,
 * DCEL_VertexImplem.hpp:190-300) and a K-ary binned-SAH BVH with W-wide leaf groups
 * (Source/EBGeometry_BVH.hpp:521-734, MeshDistanceFunctionsImplem.hpp:133-206, 386-460).
 * Faces are (v0,v1,v2) index triples; face i has metadata i. Returns NULL on bad input. */
orc_mesh* orc_mesh_create(const double* xyz, int64_t nv, const uint32_t* tri, int64_t nt,
                          int K, int W, int max_leaf_groups);
void      orc_mesh_destroy(orc_mesh* m);
int64_t   orc_mesh_num_nodes(const orc_mesh* m);

/* Per face: face normal (3), vertex pseudo-normals in facet order (9), pseudo-normals of
 * the half-edges v0->v1, v1->v2, v2->v0 (9). */
void orc_mesh_normals(const orc_mesh* m, double* fn, double* vn, double* en);

/* Signed distance and closest face (TriMeshSDF::getClosestTriangle semantics,
 * Source/EBGeometry_MeshDistanceFunctionsImplem.hpp:511-549) by BVH branch-and-bound
 * (PackedBVH::pruneTraverse scalar path, Source/EBGeometry_BVHImplem.hpp:1054-1096,
 * 1581-1605). tri may be NULL. Work is split statically over nthreads threads. */
void orc_mesh_sdf(const orc_mesh* m, const double* pts, int64_t n, double* d, uint32_t* tri,
                  int nthreads);

/* Same result by exhaustive search over every face in index order (first strict minimum
 * of |d| wins), the reference tests' brute-force cross-check (Tests/TestBVH.cpp:112-194). */
void orc_mesh_sdf_brute(const orc_mesh* m, const double* pts, int64_t n, double* d,
                        uint32_t* tri, int nthreads);

/* Traversal work of the reference's branch-and-bound per query: interior nodes whose child
 * boxes were read, leaf groups evaluated, real (non-padding) triangle lanes evaluated. */
void orc_mesh_work(const orc_mesh* m, const double* pts, int64_t n, int64_t* inner,
                   int64_t* groups, int64_t* tris);

/* One triangle, one point (TriangleSoAT::signedDistanceLane,
 * Source/EBGeometry_TriangleSoAImplem.hpp:1199-1293). v: 9 values (v0,v1,v2); fn: face
 * normal; vn: 3 vertex normals; en: 3 edge normals (v0->v1, v1->v2, v2->v0). */
double orc_triangle_sdf(const double* v, const double* fn, const double* vn, const double* en,
                        const double* p);

/* Face normal (x2-x0)x(x2-x1) scaled by the reciprocal of its length
 * (Source/EBGeometry_TriangleImplem.hpp:94-106). */
void orc_triangle_normal(const double* v, double* n);

/* Space-filling curve -------------------------------------------------------------------- */

/* SFC::computeBins<T> (Source/EBGeometry_SFCImplem.hpp:228-276): 21-bit grid bins. */
void orc_bins_f32(const float* pts, int64_t n, uint32_t* bins);
void orc_bins_f64(const double* pts, int64_t n, uint32_t* bins);
/* SFC::Morton::encode / decode (Source/EBGeometry_SFCImplem.hpp:31-82). */
void orc_morton_encode(const uint32_t* bins, int64_t n, uint64_t* codes);
void orc_morton_decode(const uint64_t* codes, int64_t n, uint32_t* bins);
/* Sort by code, stable on index (SFC::order, Source/EBGeometry_SFCImplem.hpp:279-300, with
 * the unspecified order of equal codes fixed to increasing index). */
void orc_sort_codes(const uint64_t* codes, int64_t n, uint32_t* perm);

/* CSG union ------------------------------------------------------------------------------ */

/* Union of spheres (SphereSDF, Source/EBGeometry_AnalyticDistanceFunctions.hpp:137-267),
 * axis-aligned boxes (BoxSDF, :269-423) and mesh instances
 * Translate(RotateZ(RotateY(RotateX(Scale(mesh))))) (Source/EBGeometry_TransformImplem.hpp:
 * 140-270). The value is the minimum over all primitives, which equals BVHUnionIF::value
 * (Source/EBGeometry_CSGImplem.hpp:466-501) because pruning never discards the minimum. */
orc_scene* orc_scene_create(int64_t ns, const double* sc, const double* sr, int64_t nb,
                            const double* blo, const double* bhi, int64_t nm,
                            const orc_mesh* const* meshes, const double* mt, const double* mrot,
                            const double* ms);
void       orc_scene_destroy(orc_scene* s);
void       orc_scene_value(const orc_scene* s, const double* pts, int64_t n, double* v,
                           int nthreads);

#ifdef __cplusplus
}
#endif
#endif
