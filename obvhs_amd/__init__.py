"""obvhs_amd -- MI355X (gfx950) implementation of the obvhs hot path: PLOC BVH2 construction, BVH2 -> CWBVH
conversion and CWBVH closest-hit / occlusion traversal, behind a mirror of the crate's API."""
from .api import (Bvh2, BvhBuildParams, Context, CwBvh, CwBvhBatch, DeviceArray, build_cwbvh_batch, ObvhsError, PlocBuilder, PlocSearchDistance,
                  SortPrecision, build_cwbvh_from_tris, bvh2_to_cwbvh, log2_threshold_table)
from .types import (AABB, BVH2_NODE, BVH2_NODE_SMALL, CWBVH_NODE, INVALID_ID, RAY, RAY_HIT, TRIANGLE,
                    aabbs_from_minmax, make_rays, ray_hit_none, triangles_from_vertices)

__all__ = ["CwBvhBatch", "build_cwbvh_batch", "Bvh2", "BvhBuildParams", "Context", "CwBvh", "DeviceArray", "ObvhsError", "PlocBuilder",
           "PlocSearchDistance", "SortPrecision", "build_cwbvh_from_tris", "bvh2_to_cwbvh", "log2_threshold_table",
           "AABB", "BVH2_NODE", "BVH2_NODE_SMALL", "CWBVH_NODE", "INVALID_ID", "RAY", "RAY_HIT", "TRIANGLE",
           "aabbs_from_minmax", "make_rays", "ray_hit_none", "triangles_from_vertices"]
