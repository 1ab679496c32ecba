"""Pin the body-wave misfit oracle with closed-form cases."""
import numpy as np
import pytest

from oracle import pyoracle as po

PI2_F32 = float(np.float32(6.283185))


def _data(nsrc, nrev, t, sigma):
    npair = nsrc * nrev
    tt = np.zeros((1, 3, npair))
    tt[0, 0] = t
    tt[0, 1] = sigma
    rs = np.ones((1, 2, npair), np.int32)
    return tt, rs


def test_zero_residual_gives_normalisation_only():
    nsrc, nrev = 3, 4
    pt = np.arange(12, dtype=float).reshape(1, nsrc, nrev) * 0.1
    tt, rs = _data(nsrc, nrev, pt.reshape(-1), 0.5)
    out, srd, sig, rc = po.body_misfit(pt, tt, rs, np.zeros((nsrc, 4)))
    assert rc == 0
    assert out[1] == 0 and out[2] == 0
    assert abs(out[0] - (12 * np.log(0.5) + 6.0 * np.log(PI2_F32))) < 1e-12
    np.testing.assert_array_equal(srd[0], pt.reshape(-1))


def test_single_residual_origin_time_and_missing_data():
    nsrc, nrev = 2, 2
    pt = np.zeros((1, nsrc, nrev))
    tt, rs = _data(nsrc, nrev, np.zeros(4), 2.0)
    loc = np.zeros((nsrc, 4))
    loc[1, 3] = 3.0                       # origin time of source 2 shifts its pairs
    rs[0, 0, 0] = 0                       # pair 1 has no datum: sigma -> 1, no residual
    out, srd, sig, rc = po.body_misfit(pt, tt, rs, loc, nrays=3)
    # residuals: pair2 0, pairs 3,4: 3.0 each (nrr = k + (j-1)*nrev)
    assert abs(out[2] - 18.0) < 1e-14
    assert abs(out[1] - 18.0 / 4) < 1e-14
    assert abs(out[0] - (18.0 / 8 + 3 * np.log(2.0) + 1.5 * np.log(PI2_F32))) < 1e-13
    assert sig[0, 0] == 1.0 and sig[0, 1] == 2.0


def test_sigdep_noise_model_and_fatal_zero_sigma():
    nsrc, nrev = 1, 3
    pt = np.array([[[1.0, 2.0, 4.0]]])
    tt, rs = _data(nsrc, nrev, np.array([1.0, 2.0, 4.0]), 0.0)
    out, srd, sig, rc = po.body_misfit(pt, tt, rs, np.zeros((1, 4)), sigdep=1,
                                       bnoise0=np.array([0.1]), bnoise1=np.array([0.05]))
    np.testing.assert_allclose(sig[0], [0.15, 0.25, 0.45], rtol=1e-15)
    assert rc == 0
    out, srd, sig, rc = po.body_misfit(pt, tt, rs, np.zeros((1, 4)))   # data sigma 0
    assert rc == -1


def test_interpolate3d_reproduces_linear_fields_and_nodes():
    """Interpolate3d (likelihood_body.F90:976-1007) is trilinear over the
    clamped cell: exact for a linear field inside the grid, equal to the node
    value at a node, with z in unscaled units against a scaled grid.  Outside
    the grid its |.| weights make a tent, not a linear extrapolation: for a
    constant field each axis contributes 1 - 2*dist/d beyond the edge."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 9, 7, 6
    grid = dict(xmin=-1.0, ymin=2.0, zmin=0.5 * 2.0, dx=0.5, dy=0.75, dz=0.4 * 2.0, scaling=2.0)
    xs = grid["xmin"] + grid["dx"] * np.arange(nx)
    ys = grid["ymin"] + grid["dy"] * np.arange(ny)
    zs = (grid["zmin"] + grid["dz"] * np.arange(nz)) / grid["scaling"]
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    a, b, c, d = 1.5, 0.3, -0.7, 2.1
    field = a + b * X + c * Y + d * Z
    for _ in range(200):
        p = np.array([rng.uniform(xs[0], xs[-1]), rng.uniform(ys[0], ys[-1]), rng.uniform(zs[0], zs[-1])])
        v = po.interpolate3d(grid, field, p)
        assert v == pytest.approx(a + b * p[0] + c * p[1] + d * p[2], rel=1e-12, abs=1e-12)
    rnd = rng.uniform(0, 5, (nx, ny, nz))
    for _ in range(30):
        i, j, k = rng.integers(0, nx - 1), rng.integers(0, ny - 1), rng.integers(0, nz - 1)
        assert po.interpolate3d(grid, rnd, np.array([xs[i], ys[j], zs[k]])) == pytest.approx(rnd[i, j, k], abs=1e-14)
    ones = np.ones((nx, ny, nz))
    p = np.array([xs[0] - 0.1, ys[-1] + 0.15, zs[2] + 0.01])
    want = (1 - 2 * 0.1 / grid["dx"]) * (1 - 2 * 0.15 / grid["dy"])
    assert po.interpolate3d(grid, ones, p) == pytest.approx(want, rel=1e-12)
    # the last grid plane belongs to the cell below it (index clamped to n-1)
    assert po.interpolate3d(grid, rnd, np.array([xs[-1], ys[-1], zs[-1]])) == pytest.approx(rnd[-1, -1, -1], abs=1e-14)
