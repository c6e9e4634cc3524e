"""obvhs_amd -- MI355X (gfx950) implementation of the obvhs hot path: PLOC BVH2 construction (with spatial
pre-splits, the reinsertion optimizer and partial / full rebuilds), BVH2 -> CWBVH conversion and CWBVH
closest-hit / occlusion traversal, behind a mirror of the crate's API; plus a batched form for many meshes."""
from .api import (Bvh2, BvhBuildParams, Context, CwBvh, CwBvhBatch, DeviceArray, ObvhsError, PlocBuilder,
                  PlocSearchDistance, Server, SortPrecision, build_cwbvh_batch, build_cwbvh_from_tris, bvh2_to_cwbvh,
                  compute_rebuild_path_flags, log2_threshold_table, ray_hits_none, rays_new, rt_triangles, size_class,
                  split_aabbs_preset, tune_ploc_radius,
                  tune_sort, tune_traversal)
from .types import (AABB, BVH2_NODE, BVH2_NODE_SMALL, CWBVH_NODE, INVALID_ID, RAY, RAY_HIT, RT_TRIANGLE, TRIANGLE,
                    aabbs_from_minmax, make_rays, ray_hit_none, triangles_from_vertices)

__all__ = ["Bvh2", "BvhBuildParams", "Context", "CwBvh", "CwBvhBatch", "DeviceArray", "ObvhsError", "PlocBuilder",
           "PlocSearchDistance", "Server", "SortPrecision", "build_cwbvh_batch", "build_cwbvh_from_tris",
           "bvh2_to_cwbvh", "compute_rebuild_path_flags", "log2_threshold_table", "ray_hits_none", "rays_new",
           "rt_triangles", "split_aabbs_preset",
           "size_class", "tune_sort", "tune_traversal", "tune_ploc_radius",
           "AABB", "BVH2_NODE", "BVH2_NODE_SMALL", "CWBVH_NODE", "INVALID_ID", "RAY", "RAY_HIT", "RT_TRIANGLE", "TRIANGLE",
           "aabbs_from_minmax", "make_rays", "ray_hit_none", "triangles_from_vertices"]
