"""dust_amd -- MI355X (gfx950) implementation of DUST's per-timestep hot path.

The octree FMM of the vortex-particle wake (src/fmm) and the dense AIC
assembly plus LU solve (src/linsys) run as hand-written HIP kernels in
libdust_gpu.so behind a C ABI (include/dust_gpu.h).  This package is the
Python host mirror of the reference's Fortran entry points:

    dust_amd.octree  -- initialize_octree / sort_particles / calculate_multipole /
                        apply_multipole / destroy_octree
    dust_amd.linsys  -- AIC assembly, dgetrf / dgetrs, solve_linsys, mixed solve
    dust_amd.workloads -- seeded synthetic inputs
"""
from . import _lib  # noqa: F401
from ._lib import DustError, lib  # noqa: F401

__all__ = ["DustError", "lib", "octree", "linsys", "workloads"]
