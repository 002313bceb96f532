// Host-side mesh preparation: pseudo-normals and the wide BVH, packed for upload.
#pragma once

#include "ebgpu_internal.hpp"

#include <string>
#include <vector>

namespace ebgpu {

struct PseudoNormals
{
  std::vector<double> face; // 3 per face
  std::vector<double> vert; // 3 per vertex
  std::vector<double> edge; // 3 per half-edge (3*face + i, edge v_i -> v_{i+1})
};

// Angle-weighted pseudo-normals with the reference's formulas and summation order:
// face (x2-x0)x(x2-x1)/|.| (Source/EBGeometry_DCEL_FaceImplem.hpp:138-165), edge =
// normalise(n_face + n_pairface) (DCEL_EdgeImplem.hpp:130-165), vertex = normalise(sum over
// incident faces in facet order of acos(clamp(e1.e2)) n_face) (DCEL_VertexImplem.hpp:190-300).
// Returns false (with a message) for degenerate input.
bool compute_pseudo_normals(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt,
                            PseudoNormals& out, std::string& err);

// A polygon mesh (the reference's DCEL MeshSDF input: planar simple polygons, faces given as
// offsets[nf + 1] into idx) as triangles carrying the polygon mesh's own pseudo-normals: each
// face is clipped into ears in the plane of its two non-dominant normal axes; every triangle has
// its face's normal, the triangle edges on the face boundary the face half-edge's normal and the
// edges inside the face the face normal, the vertices their angle-weighted normals over the
// polygons. The distance to a face is the minimum over its triangles, and each closest feature
// carries the normal the reference gives it, so the signed distance is the reference's. Faces
// that are degenerate, repeat a vertex, leave their plane by more than 1e-5 of their extent or
// cannot be clipped (self-intersecting) are refused.
bool polygon_mesh(const float* xyz, std::size_t nv, const uint32_t* offsets, const uint32_t* idx, std::size_t nf,
                  std::vector<uint32_t>& tri, std::vector<uint32_t>& tri_face, PseudoNormals& pn, std::string& err);

// Packed, upload-ready mesh.
struct PackedMesh
{
  int                   K = 4, leaf_size = 4;
  std::size_t           num_nodes = 0, num_leaves = 0;
  int                   stack_need = 0;
  float                 lo[3], hi[3];
  double                diag = 0;
  float                 hmax = 0; // longest triangle edge, rounded up
  std::vector<uint8_t>  nodes;  // Node<K>[num_nodes]
  std::vector<TriRec>   tris;   // leaf order
  std::vector<float>    tripos; // 12 per triangle, leaf order
  std::vector<uint32_t> feat;   // 8 per triangle, leaf order
  std::vector<float>    tsph;   // 8 per triangle (sphere, plane), leaf order, + kSphPad zero entries
  std::vector<float>    vn32, en32;
  std::vector<uint32_t> order;  // original face index of each leaf-order triangle
  std::vector<double>   vn64, en64, fn64;
};

// Binned-SAH binary build (32 bins, 3 axes, cost A_L N_L + A_R N_R as in the reference's
// SAH2WaySplit, Source/EBGeometry_BVH.hpp:521-641) with leaves of at most leaf_size
// triangles, collapsed into a K-wide tree (largest-area child expanded first).
bool build_sah_mesh(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt, int K,
                    int leaf_size, const PseudoNormals& pn, PackedMesh& out, std::string& err);

// Fill the per-triangle arrays (TriRec, tripos, feat) for triangles given in leaf order
// (order[i] = original face index) and the normal arrays. Shared by both builders.
void pack_triangles(const float* xyz, const uint32_t* tri, const std::vector<uint32_t>& order,
                    const PseudoNormals& pn, std::size_t nv, PackedMesh& out);

// GPU Morton build (morton_build.hip): the reference's bottom-up SFC partition into a balanced
// K-ary tree (Source/EBGeometry_BVHImplem.hpp:207-312) over triangle box centroids. Codes, sort
// and node boxes are computed on the device; the triangle records are packed here.
int  morton_depth(uint64_t nt, int K, int leaf_size);
bool build_morton_mesh_gpu(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt, int K, int leaf_size,
                           const PseudoNormals& pn, PackedMesh& out, std::string& err);

// Maximum traversal stack depth any query can need in a packed K-wide tree.
int compute_stack_need(const PackedMesh& m);

// A K-wide SAH BVH over arbitrary boxes (Node<K> records, DFS pre-order, root at 0); leaf
// children reference ranges of `order` (primitive indices in leaf order).
struct WideBvh
{
  std::vector<uint8_t>  nodes;
  std::size_t           num_nodes = 0, num_leaves = 0;
  int                   stack_need = 0;
  std::vector<uint32_t> order;
  float                 lo[3], hi[3];
  double                diag = 0;
};
bool build_wide_bvh(const float* blo, const float* bhi, std::size_t n, int K, int leaf_size, WideBvh& out,
                    std::string& err);

// One node of a binary BVH (the SAH builders' output before it is collapsed to K-wide).
struct BinNode
{
  float    lo[3], hi[3];
  int32_t  left, right; // -1 for a leaf
  uint32_t first, count; // range of the primitive order array
};
static_assert(sizeof(BinNode) == 40, "BinNode layout (also written by the GPU builder)");

// Collapse a binary BVH into K-wide Node<K> records (each node takes the largest-area interior
// child's children until it has K), DFS pre-order from the root; fills nodes, counts, root box,
// diagonal and stack need of out (not its order).
void collapse_binary(const std::vector<BinNode>& bn, int32_t root, int K, WideBvh& out);

// 64-bit FNV-1a over the bytes of the vertex and triangle arrays (identifies the mesh a trained
// network belongs to).
inline uint64_t mesh_hash(const float* xyz, uint64_t nv, const uint32_t* tri, uint64_t nt)
{
  uint64_t h   = 0xcbf29ce484222325ull;
  auto     mix = [&h](const void* p, std::size_t bytes) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (std::size_t i = 0; i < bytes; i++) {
      h ^= c[i];
      h *= 0x100000001b3ull;
    }
  };
  mix(&nv, sizeof(nv));
  mix(&nt, sizeof(nt));
  mix(xyz, sizeof(float) * 3 * nv);
  mix(tri, sizeof(uint32_t) * 3 * nt);
  return h;
}

} // namespace ebgpu
