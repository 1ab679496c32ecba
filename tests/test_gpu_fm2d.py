"""GPU 2-D phase traveltimes (modrays on gfx950) against the reference's golden
vectors and the oracle.

Each (period, source) runs the reference's sequential fast marching in one
lane, so every time must be identical to the oracle and to the reference's
compiled Fortran, bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "fm2d_modrays.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_golden_vectors():
    g = np.load(GOLD)
    n = int(g["n"])
    say("fm2d golden", n)
    for i in range(n):
        fom, asgr = (int(v) for v in g["opt_%d" % i])
        s, r = g["src_%d" % i], g["rec_%d" % i]
        tt = mct.modrays(g["velv_%d" % i], 0, 0, 0.5, 0.5, s[:, 0], s[:, 1], r[:, 0], r[:, 1],
                         srs=g["srs_%d" % i], fom=fom, asgr=asgr)
        np.testing.assert_array_equal(tt, g["tt_%d" % i], err_msg="golden case %d" % i)


@pytest.mark.parametrize("gd,sgdl,sgs", [((1, 1), 4, 8), ((2, 2), 4, 8), ((2, 1), 2, 3), ((1, 3), 5, 2)])
def test_multi_period_batch_matches_oracle(gd, sgdl, sgs):
    rng = np.random.default_rng(100 + 7 * gd[0] + gd[1] + sgdl + sgs)
    nvx, nvz, h = int(rng.integers(10, 30)), int(rng.integers(10, 30)), 0.6
    npd, ns, nr = 5, 6, 9
    velv = rng.uniform(2.0, 4.5, (npd, nvx + 2, nvz + 2))
    L, W = (nvx - 1) * h, (nvz - 1) * h
    ox, oz = -3.0, 1.5
    sx, sz = ox + rng.uniform(h, L - h, ns), oz + rng.uniform(h, W - h, ns)
    rx, rz = ox + rng.uniform(0, L, nr), oz + rng.uniform(0, W, nr)
    srs = (rng.random((npd, ns, nr)) < 0.7).astype(np.int32)
    srs[1, 2, :] = 0                                   # a source with no pairs at one period
    for fom in (0, 1):
        for asgr in (0, 1):
            say("fm2d batch", gd, sgdl, sgs, "fom", fom, "asgr", asgr)
            got = mct.modrays(velv, ox, oz, h, h, sx, sz, rx, rz, srs=srs, gdx=gd[0], gdz=gd[1],
                              asgr=asgr, sgdl=sgdl, sgs=sgs, fom=fom, ttime=np.full((npd, ns, nr), -1.0))
            for p in range(npd):
                want = po.fm2d(velv[p], ox, oz, h, h, sx, sz, rx, rz, srs=srs[p], gdx=gd[0], gdz=gd[1],
                               asgr=asgr, sgdl=sgdl, sgs=sgs, fom=fom, ttime=np.full((ns, nr), -1.0))
                np.testing.assert_array_equal(got[p], want, err_msg="period %d" % p)
            assert (got[srs == 0] == -1.0).all()


def test_sources_on_grid_nodes_and_receivers_on_edges():
    n, h = 25, 0.5
    rng = np.random.default_rng(3)
    velv = rng.uniform(2.5, 3.5, (n + 2, n + 2))
    L = (n - 1) * h
    sx, sz = np.array([1.0, 6.0, 3.5]), np.array([1.0, 6.0, 9.0])
    rx = np.array([0.0, L, 0.0, L, 5.25, 2.0])
    rz = np.array([0.0, 0.0, L, L, 0.0, L])
    for fom in (0, 1):
        got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom)
        want = po.fm2d(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom)
        np.testing.assert_array_equal(got, want)


def test_source_outside_the_grid_is_an_error():
    velv = np.full((12, 12), 3.0)
    with pytest.raises(mct.MCTError):
        mct.modrays(velv, 0, 0, 1.0, 1.0, np.array([50.0]), np.array([2.0]), np.array([1.0]), np.array([1.0]))
    with pytest.raises(mct.MCTError):
        mct.modrays(velv, 0, 0, 1.0, 1.0, np.array([2.0]), np.array([2.0]), np.array([1.0]), np.array([-40.0]))
