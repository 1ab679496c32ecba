"""Pin the dispersion oracle (oracle/surfdisp.c, a restatement of
surfmodes/surfdisp96.f) against the reference's own compiled Fortran and
against closed forms."""
import os

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import pyoracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden", "surfdisp96.npz")


def test_matches_reference_golden_vectors():
    g = np.load(GOLD)
    nerr = 0
    for i in range(int(g["n"])):
        thk, vp, vs, rho = g["m_%d" % i]
        iw, igr, ierr = g["opt_%d" % i]
        cp, cg, e, nd = po.surfdisp96(thk, vp, vs, rho, g["t_%d" % i], int(iw), int(igr))
        np.testing.assert_array_equal(cp, g["cp_%d" % i])
        np.testing.assert_array_equal(cg, g["cg_%d" % i])
        assert e == ierr
        nerr += ierr
    assert nerr > 0          # the fixtures include no-root failures


def test_matches_compiled_reference_live():
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(77)
    for trial in range(60):
        L = int(rng.integers(2, 15))
        thk = rng.uniform(0.2, 4.0, L); thk[-1] = 0
        vs = np.sort(rng.uniform(1.5, 4.8, L)) if trial % 2 else rng.uniform(1.5, 4.8, L)
        vp = vs * 1.73
        rho = 1.74 * vp ** 0.25
        per = np.geomspace(0.2, 8, 30)
        for iw in (1, 2):
            for igr in (0, 1):
                a = po.surfdisp96(thk, vp, vs, rho, per, iw, igr)
                b = po.surfdisp96(thk, vp, vs, rho, per, iw, igr, ref=True)
                np.testing.assert_array_equal(a[0], b[0])
                np.testing.assert_array_equal(a[1], b[1])
                assert a[2] == b[2]


def rayleigh_halfspace(alpha, beta):
    f = lambda c: (2 - c**2 / beta**2) ** 2 - 4 * np.sqrt(1 - c**2 / alpha**2) * np.sqrt(1 - c**2 / beta**2)
    return brentq(f, 0.5 * beta, beta * (1 - 1e-12), xtol=1e-14)


@pytest.mark.parametrize("beta", [1.7, 3.2, 4.4])
def test_halfspace_rayleigh_is_nondispersive_root(beta):
    """Uniform half-space: c_R solves (2-k^2)^2 = 4 sqrt(1-g^2k^2) sqrt(1-k^2),
    at every period; group velocity equals phase velocity."""
    alpha = 1.73 * beta
    thk = np.array([2.0, 0.0])
    vp = np.array([alpha, alpha])
    vs = np.array([beta, beta])
    rho = np.array([2.6, 2.6])
    per = np.geomspace(0.5, 10, 12)
    a32, b32 = float(np.float32(alpha)), float(np.float32(beta))
    cr = rayleigh_halfspace(a32, b32)
    cp, _, ierr, _ = po.surfdisp96(thk, vp, vs, rho, per, 2, 0)
    assert ierr == 0
    assert np.max(np.abs(cp - cr)) < 2e-6 * cr
    cp, cg, ierr, _ = po.surfdisp96(thk, vp, vs, rho, per, 2, 1)
    assert np.max(np.abs(cg - cr)) < 2e-3 * cr


def love_two_layer(T, h, b1, b2, r1, r2):
    """Fundamental Love mode of a layer over a half-space: the lowest c in
    (b1, b2) with mu1 s1 sin(w h s1) = mu2 s2 cos(w h s1)."""
    w = 2 * np.pi / T
    mu1, mu2 = r1 * b1 * b1, r2 * b2 * b2

    def g(c):
        s1 = np.sqrt(1 / b1**2 - 1 / c**2)
        s2 = np.sqrt(1 / c**2 - 1 / b2**2)
        return mu1 * s1 * np.sin(w * h * s1) - mu2 * s2 * np.cos(w * h * s1)
    cs = b1 + (b2 - b1) * np.linspace(1e-9, 1 - 1e-9, 20001)
    gv = g(cs)
    i = int(np.nonzero(np.sign(gv[1:]) != np.sign(gv[:-1]))[0][0])
    return brentq(g, cs[i], cs[i + 1], xtol=1e-15)


def test_love_layer_over_halfspace_closed_form():
    h, b1, b2, r1, r2 = 3.0, 2.8, 4.2, 2.5, 2.9
    thk = np.array([h, 0.0])
    vs = np.array([b1, b2])
    vp = vs * 1.73
    rho = np.array([r1, r2])
    per = np.geomspace(1, 8, 8)
    cp, _, ierr, _ = po.surfdisp96(thk, vp, vs, rho, per, 1, 0)
    assert ierr == 0
    f32 = lambda x: float(np.float32(x))
    for T, c in zip(per, cp):
        exact = love_two_layer(T, f32(h), f32(b1), f32(b2), f32(r1), f32(r2))
        assert abs(c - exact) < 3e-6 * exact, (T, c, exact)


def test_normal_dispersion_and_group_below_phase():
    rng = np.random.default_rng(4)
    for _ in range(10):
        L = int(rng.integers(3, 10))
        thk = rng.uniform(0.5, 3, L); thk[-1] = 0
        vs = np.sort(rng.uniform(2.5, 4.5, L))
        vp = 1.73 * vs
        rho = 1.74 * vp ** 0.25
        per = np.geomspace(1, 10, 15)
        for iw in (1, 2):
            cp, _, ierr, _ = po.surfdisp96(thk, vp, vs, rho, per, iw, 0)
            assert ierr == 0
            assert np.all(np.diff(cp) >= -1e-6)       # phase velocity grows with period
            cp1, cg, ierr, _ = po.surfdisp96(thk, vp, vs, rho, per, iw, 1)
            assert np.all(cg <= cp1 + 1e-4)


GOLD_W = os.path.join(os.path.dirname(__file__), "golden", "surfdisp96_water.npz")


def test_water_layer_matches_reference_golden_vectors():
    """Models under a water layer (vs = 0 on top: surfdisp96's fluid-layer
    branches, surfmodes/surfdisp96.f:1196-1215 among them), written by the
    compiled reference (tests/golden/make_golden.py surfdisp_water)."""
    g = np.load(GOLD_W)
    for i in range(int(g["n"])):
        thk, vp, vs, rho = g["m_%d" % i]
        assert vs[0] == 0 and vp[0] == 1.5
        iw, igr, ierr = g["opt_%d" % i]
        cp, cg, e, nd = po.surfdisp96(thk, vp, vs, rho, g["t_%d" % i], int(iw), int(igr))
        np.testing.assert_array_equal(cp, g["cp_%d" % i])
        np.testing.assert_array_equal(cg, g["cg_%d" % i])
        assert e == ierr


def test_water_layer_matches_compiled_reference_live():
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(78)
    for trial in range(40):
        L = int(rng.integers(2, 14))
        thk = rng.uniform(0.2, 4.0, L); thk[-1] = 0
        vs = np.sort(rng.uniform(1.0, 4.8, L)) if trial % 2 else rng.uniform(1.0, 4.8, L)
        vp = vs * 1.73
        rho = 1.74 * vp ** 0.25
        w = rng.uniform(0.05, 4.0)
        m = (np.r_[w, thk], np.r_[1.5, vp], np.r_[0.0, vs], np.r_[1.0, rho])
        per = np.geomspace(0.2, 10, 25)
        for iw in (1, 2):
            for igr in (0, 1):
                a = po.surfdisp96(*m, per, iw, igr)
                b = po.surfdisp96(*m, per, iw, igr, ref=True)
                np.testing.assert_array_equal(a[0], b[0])
                np.testing.assert_array_equal(a[1], b[1])
                assert a[2] == b[2]


def test_thin_water_layer_limits():
    """What the mathematics fixes for a water layer over a layered solid: at
    periods long against the water depth the Rayleigh phase velocity tends to
    the dry model's, and at periods short against it the fundamental mode is
    trapped at the water-solid interface, slower than sound in water (a
    Scholte wave).  A dropped water layer or a water layer read as solid fails
    one of the two."""
    thk = np.array([1.0, 2.0, 0.0]); vs = np.array([2.0, 3.0, 4.0])
    vp = vs * 1.73; rho = 1.74 * vp ** 0.25
    per_long = np.geomspace(20.0, 40.0, 5)
    dry = po.surfdisp96(thk, vp, vs, rho, per_long, 2, 0)
    wet = po.surfdisp96(np.r_[0.01, thk], np.r_[1.5, vp], np.r_[0.0, vs], np.r_[1.0, rho], per_long, 2, 0)
    assert dry[2] == 0 and wet[2] == 0
    assert np.max(np.abs(wet[0] - dry[0]) / dry[0]) < 5e-3
    per_short = np.array([0.05, 0.08])
    deep = po.surfdisp96(np.r_[3.0, thk], np.r_[1.5, vp], np.r_[0.0, vs], np.r_[1.0, rho], per_short, 2, 0)
    assert deep[2] == 0
    assert np.all(deep[0] < 1.5) and np.all(deep[0] > 1.0)
