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


@dataclass
class SurfWorkload:
    grid: Grid
    points: np.ndarray      # (ncells, 3)   RTI%points
    params: np.ndarray      # (ncells, 3)   (vp, vs, rho); only vs is used (datatype 2)
    freqs: np.ndarray       # (np,) Hz, ascending period
    src: np.ndarray         # (nsrc, 2)     dat%src(1:2,:)
    rev: np.ndarray         # (nrev, 2)     dat%rev(1:2,:)
    ttime: np.ndarray       # (np, 3, nsrc*nrev)  observed (t, sigma, -)
    raystat: np.ndarray     # (np, 2, nsrc*nrev)  int32
    raylov: int = 1
    order: int = 1
    rejected_draws: int = 0  # site sets redrawn because they failed check_model

    @property
    def nsrc(self):
        return self.src.shape[0]


def _surface_sites(rng, ncells, L, Z, vs_range):
    """Sites uniform in the box, Vs uniform in vs_range assigned in order of
    site depth, so that slower sites lie shallower."""
    pts = np.column_stack([rng.uniform(0.0, L, ncells), rng.uniform(0.0, L, ncells),
                           rng.uniform(0.0, Z, ncells)])
    vs = np.sort(rng.uniform(*vs_range, ncells))[np.argsort(np.argsort(pts[:, 2]))]
    vp = 1.73 * vs
    return pts, np.column_stack([vp, vs, 0.32 * vp + 0.77])


def top_is_slowest(grid, points, params):
    """check_model (src/likelihood_surf.F90:631-646) on the nearest-site grid,
    from a brute-force argmin of squared distances (no ties in random sites)."""
    xs = grid.xmin + grid.dx * np.arange(grid.nx)
    ys = grid.ymin + grid.dy * np.arange(grid.ny)
    zs = grid.zmin + grid.dz * np.arange(grid.nz)
    vs = params[:, 1]
    top = np.full((grid.nx, grid.ny), np.inf)
    ok = True
    for k, z in enumerate(zs):
        d = ((xs[:, None, None] - points[None, None, :, 0]) ** 2 +
             (ys[None, :, None] - points[None, None, :, 1]) ** 2) + (z - points[None, None, :, 2]) ** 2
        v = vs[np.argmin(d, axis=2)]
        if k == 0:
            top = v
        elif np.any(v < top):
            ok = False
            break
    return ok


def surface_workload(n=41, nz=21, L=20.0, Z=10.0, ncells=30, np_=10, periods=(1.0, 5.0),
                     nsrc=16, nrev=16, vs_range=(2.5, 4.5), sigma=0.05, raylov=1, order=1,
                     seed=20261202, ttime=None):
    """Surface-wave likelihood inputs: ncells Voronoi sites in an L x L x Z km
    box rasterised to n x n x nz nodes, np_ periods log-spaced in `periods`
    (freqs = 1/T, ascending T), nsrc sources and nrev receivers uniform at the
    surface at least one cell inside the model, every pair used.  Site sets
    that fail check_model are redrawn (counted in rejected_draws).  Observed
    times default to straight-ray times at 2.9 km/s plus Gaussian noise."""
    rng = np.random.default_rng(seed)
    h, hz = L / (n - 1), Z / (nz - 1)
    grid = Grid(n, n, nz, 0.0, 0.0, 0.0, h, h, hz, 1.0)
    rejected = 0
    while True:
        points, params = _surface_sites(rng, ncells, L, Z, vs_range)
        if top_is_slowest(grid, points, params):
            break
        rejected += 1
        if rejected > 100:
            raise RuntimeError("no site set passed check_model")
    T = np.geomspace(periods[0], periods[1], np_)
    freqs = 1.0 / T
    src = rng.uniform(h, L - h, (nsrc, 2))
    rev = rng.uniform(h, L - h, (nrev, 2))
    npair = nsrc * nrev
    tt = np.zeros((np_, 3, npair))
    if ttime is None:
        d = np.linalg.norm(src[:, None, :] - rev[None, :, :], axis=2).reshape(-1)
        for i in range(np_):
            tt[i, 0] = d / 2.9 + rng.normal(0.0, sigma, npair)
    else:
        tt[:, 0] = ttime
    tt[:, 1] = sigma
    # raystat(n,1) = 1 used pair, raystat(n,2) = n as read_times_3 sets it
    # (src/likelihood_settings.f90:400-401)
    rs = np.ones((np_, 2, npair), np.int32)
    rs[:, 1, :] = np.arange(1, npair + 1, dtype=np.int32)
    return SurfWorkload(grid, points, params, freqs, src, rev, tt, rs, raylov, order, rejected)


def c1_surface(seed=20261202, order=1):
    """The plain surface-wave config: 30 sites in 20x20x10 km, Vs U[2.5,4.5],
    41x41x21 grid, Rayleigh phase velocity at 10 periods 1-5 s, 16 sources
    and 16 receivers, sigma 0.05 s."""
    return surface_workload(41, 21, 20.0, 10.0, 30, 10, (1.0, 5.0), 16, 16, seed=seed, order=order)


def surface_201(seed=20261206, order=1, ncells=200, np_=20, nsrc=16, nrev=16):
    """The 201x201x61 surface-wave grid of the served mix: 100 x 100 x 30 km
    at 0.5 km, Rayleigh phase velocity at np_ periods 1-10 s."""
    return surface_workload(201, 61, 100.0, 30.0, ncells, np_, (1.0, 10.0), nsrc, nrev, seed=seed,
                            order=order)


@dataclass
class JointWorkload:
    grid: Grid
    points: np.ndarray      # (ncells, 3)   RTI%points
    params: np.ndarray      # (ncells, 3)   (vp, vs, rho), each the site's own (datatype 3)
    locations: np.ndarray   # (nsrc, 4)     RTI%locations (x, y, z, t0) of the body-wave sources
    body: dict              # BodyLikelihood-style data: rev (nrev,3), ttime (2,3,npair), raystat, nsrc
    surf: dict              # SurfLikelihood-style data: freqs, src, rev, ttime, raystat
    rejected_draws: int = 0


def _joint_sites(rng, ncells, L, Z, vs_range, pert):
    """Sites uniform in the box; Vs a linear depth gradient over vs_range plus a
    uniform perturbation of +-pert, clipped to vs_range and ranked by site depth
    (slower sites shallower); vp = Vs times U[1.65, 1.85]; rho = 0.32 vp + 0.77
    plus U[-0.05, 0.05] g/cc: three independent site parameters."""
    pts = np.column_stack([rng.uniform(0.0, L, ncells), rng.uniform(0.0, L, ncells),
                           rng.uniform(0.0, Z, ncells)])
    vs = vs_range[0] + (vs_range[1] - vs_range[0]) * pts[:, 2] / Z + rng.uniform(-pert, pert, ncells)
    vs = np.clip(vs, *vs_range)
    vs = np.sort(vs)[np.argsort(np.argsort(pts[:, 2]))]
    vp = vs * rng.uniform(1.65, 1.85, ncells)
    rho = 0.32 * vp + 0.77 + rng.uniform(-0.05, 0.05, ncells)
    return pts, np.column_stack([vp, vs, rho])


def joint_workload(n=41, nz=21, L=20.0, Z=10.0, ncells=40, nsrc=6, nrev=48, np_=6, periods=(1.0, 5.0),
                   nsta=8, vs_range=(1.5, 4.8), pert=0.3, sigma=0.05, order=0, seed=20261205,
                   check=True):
    """Joint body + surface-wave inputs (datatype 3): ncells sites in an L x L x
    Z km box on n x n x nz nodes; nsrc body-wave sources in the volume (depth
    10-90% of Z) and nrev receivers at the surface, P and S; surface-wave
    station pairs on nsta of those receivers (sources = receivers = the
    stations, as the reference's example surface data) at np_ periods
    log-spaced over `periods`.  Observed times: straight rays at 6 / 3.5 km/s
    (body) and 3 km/s (surface) plus Gaussian noise.  check: redraw site sets
    that fail check_model (brute-force nearest site on the host)."""
    rng = np.random.default_rng(seed)
    h, hz = L / (n - 1), Z / (nz - 1)
    grid = Grid(n, n, nz, 0.0, 0.0, 0.0, h, h, hz, 1.0)
    rejected = 0
    while True:
        points, params = _joint_sites(rng, ncells, L, Z, vs_range, pert)
        if not check or top_is_slowest(grid, points, params):
            break
        rejected += 1
        if rejected > 100:
            raise RuntimeError("no site set passed check_model")
    loc = np.zeros((nsrc, 4))
    loc[:, 0:2] = rng.uniform(h, L - h, (nsrc, 2))
    loc[:, 2] = rng.uniform(0.1 * Z, 0.9 * Z, nsrc)
    rev = np.zeros((nrev, 3))
    rev[:, 0:2] = rng.uniform(h, L - h, (nrev, 2))
    d = np.linalg.norm(loc[:, None, 0:3] - rev[None, :, :], axis=2).reshape(-1)
    npair = nsrc * nrev
    btt = np.zeros((2, 3, npair))
    for i, v in enumerate((6.0, 3.5)):
        btt[i, 0] = d / v + rng.normal(0.0, sigma, npair)
    btt[:, 1] = sigma
    brs = np.ones((2, 2, npair), np.int32)
    sta = rev[:nsta, 0:2].copy()
    T = np.geomspace(periods[0], periods[1], np_)
    sp = nsta * nsta
    ds = np.linalg.norm(sta[:, None, :] - sta[None, :, :], axis=2).reshape(-1)
    stt = np.zeros((np_, 3, sp))
    srs = np.zeros((np_, 2, sp), np.int32)
    used = ds > 0.0                         # a station is not paired with itself
    for i in range(np_):
        stt[i, 0] = ds / 3.0 + rng.normal(0.0, sigma, sp)
        srs[i, 0] = used
        srs[i, 1] = np.where(used, np.arange(1, sp + 1), 0)
    stt[:, 1] = sigma
    body = dict(rev=rev, ttime=btt, raystat=brs, nsrc=nsrc, order=order)
    surf = dict(freqs=1.0 / T, src=sta, rev=sta.copy(), ttime=stt, raystat=srs, order=order)
    return JointWorkload(grid, points, params, loc, body, surf, rejected)


def c5_joint(seed=20261205, order=0, check=False):
    """The joint config: 201x201x81 (100 x 100 x 40 km at 0.5 km), 600 sites, Vs
    1.5-4.8 km/s with a linear depth gradient plus +-0.3 km/s, 400 sources in
    the volume and 2000 surface receivers (P and S), surface waves at 30
    periods 0.2-8 s on 64 of the receivers as stations (64 x 63 pairs)."""
    return joint_workload(201, 81, 100.0, 40.0, 600, 400, 2000, 30, (0.2, 8.0), 64, order=order, seed=seed,
                          check=check)
