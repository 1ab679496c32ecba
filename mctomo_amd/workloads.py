"""Seeded synthetic workloads for the body-wave and surface-wave likelihood.

The counterpart of the reference's forward-modelling harness
(src/forward_modelling.f90:221-544): it builds the inputs a likelihood call
takes -- grid, Voronoi sites and their parameters, sources, receivers and
observed data -- from a seed, with no files.
"""
from dataclasses import dataclass

import numpy as np

from .api import Grid


@dataclass
class BodyWorkload:
    grid: Grid
    points: np.ndarray      # (ncells, 3)   RTI%points, z scaled
    params: np.ndarray      # (ncells, 3)   (vp, vs, rho)
    locations: np.ndarray   # (nsrc, 4)     RTI%locations (x, y, z, t0)
    rev: np.ndarray         # (nrev, 3)     dat%rev
    ttime: np.ndarray       # (np, 3, nsrc*nrev)  observed (t, sigma, -)
    raystat: np.ndarray     # (np, 2, nsrc*nrev)  int32
    datatype: int = 0
    order: int = 0

    @property
    def nsrc(self):
        return self.locations.shape[0]


def body_workload(n=101, h=0.2, ncells=100, nsrc=50, nrev=400, datatype=0, order=0,
                  vp_range=(4.0, 7.0), src_depth=(2.0, 18.0), sigma=0.05, seed=20261203,
                  ttime=None):
    """3-D body-wave eikonal drop-in: Voronoi Vp uniform in vp_range on an n^3
    grid of spacing h, sources uniform at depth src_depth, receivers at z=0.
    Observed times default to straight-ray times at the mean velocity plus
    Gaussian noise (the caller may pass exact ones)."""
    rng = np.random.default_rng(seed)
    L = h * (n - 1)
    grid = Grid(n, n, n, 0.0, 0.0, 0.0, h, h, h, 1.0)
    points = rng.uniform(0.0, L, (ncells, 3))
    vp = rng.uniform(*vp_range, ncells)
    params = np.stack([vp, vp / 1.73, 0.32 * vp + 0.77], axis=1)
    loc = np.zeros((nsrc, 4))
    loc[:, 0:2] = rng.uniform(0.0, L, (nsrc, 2))
    loc[:, 2] = rng.uniform(*src_depth, nsrc)
    rev = np.zeros((nrev, 3))
    rev[:, 0:2] = rng.uniform(0.0, L, (nrev, 2))
    npp = 1 if datatype == 0 else 2
    npair = nsrc * nrev
    tt = np.zeros((npp, 3, npair))
    if ttime is None:
        d = np.linalg.norm(loc[:, None, 0:3] - rev[None, :, :], axis=2).reshape(-1)
        vmean = np.mean(vp_range)
        for i in range(npp):
            v = vmean if i == 0 else vmean / 1.73
            tt[i, 0] = d / v + rng.normal(0.0, sigma, npair)
    else:
        tt[:, 0] = ttime
    tt[:, 1] = sigma
    rs = np.ones((npp, 2, npair), np.int32)
    return BodyWorkload(grid, points, params, loc, rev, tt, rs, datatype, order)


def c2_body(seed=20261203, order=0):
    """The 3-D body-wave config: 100 sites, Vp U[4,7] km/s, 101^3 grid at
    0.2 km, 50 sources at depth 2-18 km, 400 surface receivers, P only."""
    return body_workload(101, 0.2, 100, 50, 400, 0, order, seed=seed)
