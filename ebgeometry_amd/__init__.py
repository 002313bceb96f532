"""ebgeometry_amd -- MI355X-native batched signed-distance queries.

A from-scratch gfx950 implementation of the reference library's signed-distance query path
(packed-BVH traversal over triangle meshes and CSG primitives, the point-triangle
closest-feature distance and the pseudo-normal sign test), behind the reference's
ImplicitFunction / SignedDistanceFunction / TriMeshSDF interface. Compute runs in
libebgpu.so (hand-written HIP kernels + C ABI, include/ebgpu.h); this package is the host
mirror of the reference API over that ABI.
"""
from .sdf import (Build, ImplicitFunction, MeshSDF, SignedDistanceFunction, TriMeshSDF, grid_points, morton_codes,
                  morton_sort, polygons_to_triangles, pseudo_normals)
from .batch import BatchPlan, MeshBatch
from .cloud import Hit, PointCloudBVH, PointCloudHashGrid
from .csg import (AABB, Annular, AnnularIF, Blur, BlurIF, BoxSDF, BVHSmoothUnion, BVHSmoothUnionIF, BVHUnion,
                  BVHUnionIF, CapsuleSDF, Complement, ComplementIF, ConeSDF, CylinderSDF, Difference, DifferenceIF,
                  Elongate, ElongateIF, ExpMin, FiniteRepetition, FiniteRepetitionIF, InfiniteConeSDF,
                  InfiniteCylinderSDF, Intersection, IntersectionIF, Mollify, MollifyIF, Offset, OffsetIF, PlaneSDF,
                  Reflect, ReflectIF, Rotate, RotateIF, RoundedBoxSDF, RoundedCylinderSDF, Scale, ScaleIF,
                  SmoothDifference, SmoothDifferenceIF, SmoothIntersection, SmoothIntersectionIF, SmoothMin,
                  SmoothUnion, SmoothUnionIF, SphereSDF, TorusSDF, Translate, TranslateIF, Union, UnionIF)

__all__ = ["Build", "ImplicitFunction", "SignedDistanceFunction", "TriMeshSDF", "grid_points", "morton_codes",
           "morton_sort",
           "pseudo_normals", "AABB", "BoxSDF", "BVHUnion", "BVHUnionIF", "Rotate", "RotateIF", "Scale", "ScaleIF",
           "SphereSDF", "Translate", "TranslateIF", "MeshBatch", "BatchPlan", "Annular", "AnnularIF", "Blur",
           "BlurIF", "BVHSmoothUnion", "BVHSmoothUnionIF", "CapsuleSDF", "Complement", "ComplementIF", "ConeSDF",
           "CylinderSDF", "Difference", "DifferenceIF", "Elongate", "ElongateIF", "ExpMin", "FiniteRepetition",
           "FiniteRepetitionIF", "InfiniteConeSDF", "InfiniteCylinderSDF", "Intersection", "IntersectionIF",
           "Mollify", "MollifyIF", "Offset", "OffsetIF", "PlaneSDF", "Reflect", "ReflectIF", "RoundedBoxSDF",
           "RoundedCylinderSDF", "SmoothDifference", "SmoothDifferenceIF", "SmoothIntersection",
           "SmoothIntersectionIF", "SmoothMin", "SmoothUnion", "SmoothUnionIF", "TorusSDF", "Union", "UnionIF",
           "Hit", "PointCloudBVH", "PointCloudHashGrid", "MeshSDF", "polygons_to_triangles"]
