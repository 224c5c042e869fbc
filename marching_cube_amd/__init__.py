"""marching_cube_amd -- gfx950-native TSDF -> mesh path (integration, fused
marching-cubes extraction, normals, differentiable backward, batching)
behind the reference's host interface.  Kernels live in libmcx.so (C ABI,
include/mcx.h); this package is the host-side mirror and torch plumbing."""
__version__ = "0.1.0"
