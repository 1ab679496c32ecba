"""Pin the surface-wave likelihood oracle (oracle/surf.c) stage by stage.

Each stage it composes is pinned on its own against the reference's compiled
code (nearest site, surfdisp96, modrays).  Here the glue between them --
vs2vp/vp2rho, check_model, convert_to_layer, the padded velocity grid, the
misfit -- is checked by independent constructions: layers from a run-length
encoding of each column, the padded grid from numpy edge padding, the misfit
from the body-wave misfit (pinned separately) with zero origin times.
"""
import numpy as np
import pytest

from mctomo_amd import workloads as W
from oracle import pyoracle as po


def gdict(g):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin, dx=g.dx,
                dy=g.dy, dz=g.dz, scaling=g.scaling)


@pytest.fixture(scope="module")
def small():
    w = W.surface_workload(n=17, nz=11, L=8.0, Z=5.0, ncells=12, np_=6, periods=(1.0, 4.0),
                           nsrc=5, nrev=7, seed=11)
    res = po.surf_likelihood(gdict(w.grid), w.points, w.params, w.freqs, w.src, w.rev, w.ttime,
                             w.raystat, order=1)
    return w, res


def column_layers_rle(vs_col, dz):
    """Layers of one column: runs of equal vs; the last run is the half-space."""
    starts = np.flatnonzero(np.r_[True, vs_col[1:] != vs_col[:-1]])
    lens = np.diff(np.r_[starts, vs_col.shape[0]])
    vs = vs_col[starts]
    thick = lens * dz
    thick[-1] = 0.0
    vp = vs * np.float64(np.float32(1.730))
    rho = np.float64(np.float32(1.74)) * vp ** 0.25
    return thick, vp, vs, rho


def test_phase_velocities_are_surfdisp96_of_each_column(small):
    w, (rc, out, pt, sd, sg, pv, sid) = small
    assert rc == 0
    _, _, vs, _ = po.kdtree_to_grid(gdict(w.grid), w.points, w.params)
    T = 1.0 / w.freqs
    for ix in range(w.grid.nx):
        for iy in range(w.grid.ny):
            th, a, b, r = column_layers_rle(vs[ix, iy], w.grid.dz)
            cp, cg, e, _ = po.surfdisp96(th, a, b, r, T, 2, 0)
            assert e == 0
            np.testing.assert_array_equal(pv[ix, iy], cp)


def test_phase_times_are_modrays_on_edge_padded_grids(small):
    w, (rc, out, pt, sd, sg, pv, sid) = small
    g = w.grid
    for p in range(w.freqs.shape[0]):
        velv = np.pad(pv[:, :, p], 1, mode="edge")
        tt = po.fm2d(velv, g.xmin, g.ymin, g.dx, g.dy, w.src[:, 0], w.src[:, 1], w.rev[:, 0],
                     w.rev[:, 1], srs=w.raystat[p, 0], fom=1)
        np.testing.assert_array_equal(pt[p], tt)


def test_misfit_matches_the_body_misfit_with_zero_origin_times(small):
    w, (rc, out, pt, sd, sg, pv, sid) = small
    loc = np.zeros((w.nsrc, 4))
    o2, sd2, sg2, rc2 = po.body_misfit(pt, w.ttime, w.raystat, loc)
    assert rc2 == 0
    np.testing.assert_array_equal(out, o2)
    np.testing.assert_array_equal(sd, pt.reshape(sd.shape))
    np.testing.assert_array_equal(sg, sg2)


def test_sigma_depends_on_distance_when_sigdep(small):
    w, _ = small
    s0, s1 = np.linspace(0.01, 0.02, 6), np.full(6, 0.03)
    rc, out, pt, sd, sg, pv, sid = po.surf_likelihood(
        gdict(w.grid), w.points, w.params, w.freqs, w.src, w.rev, w.ttime, w.raystat,
        sigdep=1, snoise0=s0, snoise1=s1)
    assert rc == 0
    np.testing.assert_array_equal(sg, s0[:, None] * pt.reshape(6, -1) + s1[:, None])
    r = pt.reshape(6, -1) - w.ttime[:, 0]
    want = np.sum(r ** 2 / (2 * sg ** 2)) + np.sum(np.log(sg)) + w.raystat[:, 0].sum() / 2.0 * np.log(np.float64(np.float32(6.283185)))
    assert abs(out[0] - want) <= 1e-9 * abs(want)


def test_unused_pairs_take_unit_sigma_and_no_residual(small):
    w, _ = small
    rs = w.raystat.copy()
    rs[:, 0, ::3] = 0
    rc, out, pt, sd, sg, pv, sid = po.surf_likelihood(
        gdict(w.grid), w.points, w.params, w.freqs, w.src, w.rev, w.ttime, rs)
    assert rc == 0
    assert (sg[:, ::3] == 1.0).all()
    used = rs[:, 0] == 1
    r = (pt.reshape(used.shape) - w.ttime[:, 0])[used]
    assert abs(out[2] - np.sum(r ** 2)) <= 1e-12 * out[2]


def test_slow_layer_below_the_top_is_rejected(small):
    w, _ = small
    prm = w.params.copy()
    deepest = np.argmax(w.points[:, 2])
    prm[deepest, 1] = prm[:, 1].min() - 0.3      # slower than every top node
    rc, out, *_ = po.surf_likelihood(gdict(w.grid), w.points, prm, w.freqs, w.src, w.rev,
                                     w.ttime, w.raystat)
    assert rc == 1


def test_vs_only_vp_and_rho_of_the_sites_are_ignored(small):
    w, (rc, out, *_) = small
    prm = w.params.copy()
    prm[:, 0] *= 1.1
    prm[:, 2] += 0.5
    rc2, out2, *_ = po.surf_likelihood(gdict(w.grid), w.points, prm, w.freqs, w.src, w.rev,
                                       w.ttime, w.raystat)
    assert rc2 == 0
    np.testing.assert_array_equal(out, out2)


def test_reference_stages_equal_the_oracle_live():
    """surfdisp96 per column and modrays per period from the reference's own
    compiled Fortran (oracle/_ref, OpenMP as its OMP build), composed as
    surf_likelihood composes them, equal the oracle's phase velocities and
    phase times bit for bit (orders 0 and 1)."""
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    for order in (0, 1):
        w = W.surface_workload(n=41, nz=21, ncells=30, np_=10, nsrc=16, nrev=16, seed=20261202, order=order)
        g = gdict(w.grid)
        nbad, pv, pt = po.ref_surf_forward(g, w.points, w.params, w.freqs, w.src, w.rev, w.raystat,
                                           order=order, nthreads=4)
        res = po.surf_likelihood(g, w.points, w.params, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=order)
        assert nbad == 0 and res[0] == 0
        np.testing.assert_array_equal(pv, res[5])
        np.testing.assert_array_equal(pt, res[2])
