"""GPU dispersion (surfdisp96 on gfx950) against the oracle and the reference's
golden vectors.

The golden vectors (the reference's Fortran, compiled) must match exactly.
The period equations take exp, sin and cos from libm_glibc.h, which returns
glibc's numbers (the oracle's and the reference's), so on random layered
models the roots follow the reference's search path and the phase and group
velocities meet the 1e-6 km/s gate; the tests also count the models that are
bit-identical.
"""
import os
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "surfdisp96.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_golden_vectors():
    g = np.load(GOLD)
    n = int(g["n"])
    groups = {}
    for i in range(n):
        iw, igr, _ = (int(v) for v in g["opt_%d" % i])
        key = (iw, igr, len(g["t_%d" % i]))
        groups.setdefault(key, []).append(i)
    worst = 0.0
    for (iw, igr, npp), idx in groups.items():
        say("golden", iw, igr, npp, len(idx))
        models = [tuple(g["m_%d" % i]) for i in idx]
        ph, gr, ie = mct.surfdisp96_batch(models, g["t_%d" % idx[0]], iw, igr)
        for r, i in enumerate(idx):
            assert ie[r] == int(g["opt_%d" % i][2])
            worst = max(worst, np.max(np.abs(ph[r] - g["cp_%d" % i])),
                        np.max(np.abs(gr[r] - g["cg_%d" % i])))
    say("golden worst abs diff km/s", worst)
    assert worst <= 1e-6


@pytest.mark.parametrize("iw,igr", [(2, 0), (1, 0), (2, 1), (1, 1)])
def test_random_models_vs_oracle(iw, igr):
    rng = np.random.default_rng(10 + iw * 2 + igr)
    models = []
    for t in range(300):
        L = int(rng.integers(2, 25))
        thk = rng.uniform(0.2, 3.0, L); thk[-1] = 0
        vs = rng.uniform(1.5, 4.8, L)
        if t % 2 == 0:
            vs = np.sort(vs)
        vp = 1.73 * vs
        models.append((thk, vp, vs, 1.74 * vp ** 0.25))
    per = np.geomspace(0.5, 10, 20)
    say("random", iw, igr)
    ph, gr, ie = mct.surfdisp96_batch(models, per, iw, igr)
    nexact = 0
    wp = wg = 0.0
    for r, m in enumerate(models):
        cp, cg, e, _ = po.surfdisp96(*m, per, iw, igr)
        assert ie[r] == e, (r, ie[r], e)
        dp = np.max(np.abs(ph[r] - cp))
        dg = np.max(np.abs(gr[r] - cg)) if igr else 0.0
        wp, wg = max(wp, dp), max(wg, dg)
        nexact += dp == 0 and dg == 0
    say("random", iw, igr, "exact models", nexact, "/", len(models), "worst phase", wp, "group", wg)
    assert wp <= 1e-6
    assert wg <= 1e-6
    assert nexact >= 0.99 * len(models)


@pytest.mark.parametrize("mode", [3, 1, 2])
@pytest.mark.parametrize("iw,igr", [(2, 0), (1, 1)])
def test_random_columns_through_likelihood_kernels(mode, iw, igr):
    """Random layered models written as Vs columns (layer thicknesses whole
    multiples of dz) through the likelihood's own dispersion kernels:
    check_model, convert_to_layer and surfdisp96 per column.  mode 3 is the
    column queue the whole-grid likelihood uses (layer-count order, identical
    columns solved once: every fifth column repeats an earlier one)."""
    rng = np.random.default_rng(200 + 10 * mode + iw * 2 + igr)
    dz, nz = 0.35, 96
    ncol = 400
    cols = np.zeros((ncol, nz))
    POISSON = float(np.float32(1.73))
    for c in range(ncol):
        if c % 5 == 4:
            cols[c] = cols[int(rng.integers(0, c))]
            continue
        L = int(rng.integers(2, 12))
        n = rng.integers(1, 8, L)
        vs = rng.uniform(1.5, 4.8, L)
        if c % 2 == 0:
            vs = np.sort(vs)
        v = np.repeat(vs, n)
        cols[c, :min(nz, v.size)] = v[:nz]
        cols[c, v.size:] = vs[-1]
    per = np.geomspace(0.5, 10, 20)
    say("columns", mode, iw, igr)
    ph, gr, ie, bad = mct.dispersion_columns(cols, dz, per, iw, igr, mode=mode)
    nexact = 0
    wp = wg = 0.0
    for c in range(ncol):
        # convert_to_layer on the host: a new layer where Vs changes by more than EPS
        v = cols[c]
        edges = [0] + [k for k in range(1, nz) if abs(v[k] - v[k - 1]) > float(np.float32(1e-10))]
        thk = np.array([(edges[i + 1] - edges[i]) * dz for i in range(len(edges) - 1)] + [0.0])
        vsl = v[edges]
        vp = vsl * POISSON
        rho = float(np.float32(1.74)) * vp ** 0.25
        cp, cg, e, _ = po.surfdisp96(thk, vp, vsl, rho, per, iw, igr)
        assert ie[c] == e, (c, ie[c], e)
        dp = np.max(np.abs(ph[c] - cp))
        dg = np.max(np.abs(gr[c] - cg)) if igr else 0.0
        wp, wg = max(wp, dp), max(wg, dg)
        nexact += dp == 0 and dg == 0
    say("columns", mode, iw, igr, "exact", nexact, "/", ncol, "worst phase", wp, "group", wg, "bad_model", bad)
    assert wp <= 1e-6 and wg <= 1e-6
    assert nexact >= 0.99 * ncol
    assert bad == 1      # the unsorted columns fail check_model


@pytest.mark.parametrize("mode,water", [(3, 1.3), (1, 1.3), (2, 1.3), (3, 1e-12)])
@pytest.mark.parametrize("iw,igr", [(2, 0), (1, 1), (2, 1)])
def test_columns_under_a_water_layer(mode, water, iw, igr):
    """T_GRID%waterDepth > EPS puts a water layer (vp 1.5, vs 0, rho 1,
    thickness waterDepth, not scaled) on top of every column before surfdisp96
    (src/likelihood_surf.F90:32-33,546-551); 0 < waterDepth <= EPS instead sets
    the first layer's thickness to it (:616-617).  The likelihood's kernels
    against the oracle's surfdisp96 (pinned to the compiled reference on
    water models) on the layers that rule gives, bit for bit; with scaling 2
    the solid layers' thicknesses are halved and the water's is not."""
    rng = np.random.default_rng(900 + 10 * mode + iw * 2 + igr + (water < 1))
    dz, nz, scaling = 0.35, 60, 2.0
    ncol = 160
    cols = np.zeros((ncol, nz))
    POISSON = float(np.float32(1.73))
    EPS = float(np.float32(1e-10))
    for c in range(ncol):
        if c % 5 == 4:
            cols[c] = cols[int(rng.integers(0, c))]
            continue
        L = int(rng.integers(1, 10))
        n = rng.integers(1, 8, L)
        vs = rng.uniform(1.5, 4.8, L)
        if c % 3:
            vs = np.sort(vs)
        v = np.repeat(vs, n)
        cols[c, :min(nz, v.size)] = v[:nz]
        cols[c, v.size:] = vs[-1]
    per = np.geomspace(0.3, 10, 16)
    say("water columns", mode, water, iw, igr)
    ph, gr, ie, bad = mct.dispersion_columns(cols, dz, per, iw, igr, mode=mode, scaling=scaling, water=water)
    nexact = 0
    for c in range(ncol):
        v = cols[c]
        edges = [0] + [k for k in range(1, nz) if abs(v[k] - v[k - 1]) > EPS]
        thk = np.array([(edges[i + 1] - edges[i]) * dz / scaling for i in range(len(edges) - 1)] + [0.0])
        vsl = v[edges]
        vp = vsl * POISSON
        rho = float(np.float32(1.74)) * vp ** 0.25
        if water > EPS:
            thk, vp, vsl, rho = np.r_[water, thk], np.r_[1.5, vp], np.r_[0.0, vsl], np.r_[1.0, rho]
        else:
            thk[0] = water
        cp, cg, e, _ = po.surfdisp96(thk, vp, vsl, rho, per, iw, igr)
        assert ie[c] == e, (c, ie[c], e)
        same = np.array_equal(ph[c], cp) and (not igr or np.array_equal(gr[c], cg))
        nexact += same
    say("water columns", mode, water, iw, igr, "exact", nexact, "/", ncol, "bad_model", bad)
    assert nexact == ncol
