// GPU mesh build (gpu_build.hip): pseudo-normals, the binned-SAH BVH and the packed triangle
// records computed on the device; the host only collapses the binary tree (O(nodes)).
#pragma once

#include "host_common.hpp"
#include "mesh_build.hpp"

#include <string>

namespace ebgpu {

struct GpuBuiltMesh
{
  // K, leaf_size, node and leaf counts, root box, diagonal, hmax, stack need, the K-wide nodes
  // and the leaf order (host copies); the per-triangle and per-vertex arrays are on the device:
  PackedMesh meta;
  DevBuf     tris, tripos, tsph, feat, vn32, en32, vn64, en64, fn64, area_cdf;
  double     ms_normals = 0, ms_sah = 0, ms_pack = 0;
};

// Build mesh (xyz, nv, tri, nt) on the device with branching K and leaves of at most leaf_size
// triangles. pn: pseudo-normals computed on the host (polygon meshes), or null to compute the
// triangle mesh's own on the device. Returns EBG_OK, EBG_INVALID_ARG (degenerate input, err says
// which) or EBG_HIP_ERROR / EBG_OOM.
ebg_status build_mesh_gpu_sah(const float* xyz, std::size_t nv, const uint32_t* tri, std::size_t nt, int K,
                              int leaf_size, const PseudoNormals* pn, GpuBuiltMesh& out, std::string& err);

} // namespace ebgpu
