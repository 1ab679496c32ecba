"""mctomo_amd -- MI355X (gfx950) likelihood forward model for MCTomo.

Python host layer over the C ABI of libmctomo_hip.so (include/mctomo_hip.h).
It mirrors the reference's call surface for the hot path:

* ``fast_marching_3d``  -- the C ABI ``fastMarching3D``
  (fastmarching/fastMarching.cpp:26-129, Fortran wrapper
  src/fastMarching_wrapper.f90:51-74);
* ``kdtree_to_grid``    -- ``kdtree_to_grid`` (src/mcmc_loc2.f90:2002-2082);
* ``BodyLikelihood``    -- ``like_setup`` / ``likelihood`` for body waves
  (src/likelihood.f90:31-93, src/likelihood_body.F90:151-337).

Arrays use the reference's Fortran storage seen from C: a grid array
``(nz,ny,nx)`` in Fortran is a C-contiguous numpy array of shape
``(nx,ny,nz)``; ``points(3,ncells)`` is ``(ncells,3)``.

The library must be built (``python -m mctomo_amd._build``).  Nothing here
falls back to CPU code: a missing library raises.
"""
from .api import (  # noqa: F401
    MCTError, Grid, DeviceArray, lib, lib_path, version, fast_marching_3d, kdtree_to_grid, surfdisp96_batch, dispersion_columns, modrays, modrays_last_handover,
    BodyLikelihood, SurfLikelihood, SurfServer, JointLikelihood, REJECT_MODEL, REJECT_RAY,
    read_sources, read_times, read_data, write_sources, write_times,
)
