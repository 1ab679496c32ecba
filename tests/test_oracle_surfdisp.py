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
