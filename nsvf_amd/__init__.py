"""MI355X-native sparse-voxel rendering (the NSVF hot path) for gfx950.

Submodules:
  _ext      the reference's native op names over the C ABI (aabb_intersect,
            svo_intersect, uniform_ray_sampling, inverse_cdf_sampling,
            build_octree)
  clib      the reference's Python op forms (aabb_ray_intersect, ...)
  geometry  voxel/octree helpers
  scene     synthetic scenes of the measured configurations
"""
__version__ = "0.1.0"
