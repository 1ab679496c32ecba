"""GPU 2-D phase traveltimes (modrays on gfx950) against the reference's golden
vectors and the oracle.

method 1 runs the reference's sequential fast marching in one lane per
(period, source): every time must equal the oracle's and the reference's
compiled Fortran's, bit for bit.  method 0 runs the heap loop only near the
source and sweeps the rest to the fixed point "each node's update uses the
neighbours with smaller times".  For the first-order update (fom 0) that is
the heap loop's answer.  The mixed-order update (fom 1) can raise a node's key
when a neighbour freezes (fouds2 overwrites, updtree only sifts up), and then
the heap loop freezes nodes out of time order, which no fixed point follows.
So at fom 1 method 0 runs the heap loop itself, one wave per problem; with
MCT_FM2D_ORDER1=check it sweeps instead, and k_fm2d_verify checks every
converged field against the heap loop's freeze order and hands the problems
that fail to the sequential solvers.  Both orders are held to 1e-9 relative
(identical in practice).
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


def agree(got, want, method, msg="", tol=1e-9):
    if method == 1:
        np.testing.assert_array_equal(got, want, err_msg=msg)
        return 1.0
    m = np.isfinite(want) & (want > 0)
    rel = np.abs(got[m] - want[m]) / want[m]
    assert rel.max() <= tol, (msg, rel.max())
    assert np.array_equal(got[~m], want[~m]), msg
    return float(np.mean(got == want))


@pytest.mark.parametrize("method", [1, 0])
def test_golden_vectors(method):
    g = np.load(GOLD)
    n = int(g["n"])
    say("fm2d golden", n, "method", method)
    ex = []
    for i in range(n):
        fom, asgr = (int(v) for v in g["opt_%d" % i])
        s, r = g["src_%d" % i], g["rec_%d" % i]
        tt = mct.modrays(g["velv_%d" % i], 0, 0, 0.5, 0.5, s[:, 0], s[:, 1], r[:, 0], r[:, 1],
                         srs=g["srs_%d" % i], fom=fom, asgr=asgr, method=method)
        ex.append(agree(tt, g["tt_%d" % i], method, "golden case %d" % i, tol=1e-9))
    say("  identical fraction", np.mean(ex))


@pytest.mark.parametrize("method", [1, 0])
@pytest.mark.parametrize("gd,sgdl,sgs", [((1, 1), 4, 8), ((2, 2), 4, 8), ((2, 1), 2, 3), ((1, 3), 5, 2)])
def test_multi_period_batch_matches_oracle(gd, sgdl, sgs, method):
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
            say("fm2d batch", gd, sgdl, sgs, "fom", fom, "asgr", asgr, "method", method)
            got = mct.modrays(velv, ox, oz, h, h, sx, sz, rx, rz, srs=srs, gdx=gd[0], gdz=gd[1],
                              asgr=asgr, sgdl=sgdl, sgs=sgs, fom=fom, ttime=np.full((npd, ns, nr), -1.0),
                              method=method)
            for p in range(npd):
                want = po.fm2d(velv[p], ox, oz, h, h, sx, sz, rx, rz, srs=srs[p], gdx=gd[0], gdz=gd[1],
                               asgr=asgr, sgdl=sgdl, sgs=sgs, fom=fom, ttime=np.full((ns, nr), -1.0))
                agree(got[p], want, method, "period %d" % p, tol=1e-9)
            assert (got[srs == 0] == -1.0).all()


@pytest.mark.parametrize("method", [1, 0])
def test_sources_on_grid_nodes_and_receivers_on_edges(method):
    n, h = 25, 0.5
    rng = np.random.default_rng(3)
    velv = rng.uniform(2.5, 3.5, (n + 2, n + 2))
    L = (n - 1) * h
    sx, sz = np.array([1.0, 6.0, 3.5]), np.array([1.0, 6.0, 9.0])
    rx = np.array([0.0, L, 0.0, L, 5.25, 2.0])
    rz = np.array([0.0, 0.0, L, L, 0.0, L])
    for fom in (0, 1):
        got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom, method=method)
        want = po.fm2d(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom)
        agree(got, want, method, tol=1e-9)


@pytest.mark.parametrize("method", [1, 0])
def test_source_outside_the_grid_is_an_error(method):
    velv = np.full((12, 12), 3.0)
    with pytest.raises(mct.MCTError):
        mct.modrays(velv, 0, 0, 1.0, 1.0, np.array([50.0]), np.array([2.0]), np.array([1.0]), np.array([1.0]),
                    method=method)
    with pytest.raises(mct.MCTError):
        mct.modrays(velv, 0, 0, 1.0, 1.0, np.array([2.0]), np.array([2.0]), np.array([1.0]), np.array([-40.0]),
                    method=method)


@pytest.mark.parametrize("n,blk", [(101, 3), (101, 12), (201, 0)])
def test_parallel_path_on_large_and_contrasted_grids(n, blk):
    """Large grids and a 2.0/4.5 km/s checkerboard (long heap detours near the
    source; problems that leave the LDS neighbourhood take the sequential path)."""
    rng = np.random.default_rng(n + blk)
    h = 0.5
    L = (n - 1) * h
    if blk:
        ii = np.arange(n + 2)
        velv = np.where(((ii[:, None] // blk) + (ii[None, :] // blk)) % 2 == 0, 2.0, 4.5)
    else:
        from scipy.ndimage import zoom
        base = rng.uniform(2.6, 3.4, (n // 10 + 2, n // 10 + 2))
        velv = zoom(base, (n + 2) / base.shape[0], order=1)[:n + 2, :n + 2]
    ns, nr = 6, 12
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, L - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, L, nr)
    for fom in (0, 1):
        say("fm2d large", n, blk, "fom", fom)
        got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom, method=0)
        want = po.fm2d(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom)
        say("  identical fraction", agree(got, want, 0, tol=1e-9), "handed over", mct.modrays_last_handover())


@pytest.mark.parametrize("n,blk", [(37, 0), (64, 0), (65, 4), (130, 0), (201, 6)])
def test_strip_sweep_matches_anti_diagonal_sweep(n, blk, monkeypatch):
    """The strip-pipelined sweep visits nodes in another order that reads the
    same values as the anti-diagonal Gauss-Seidel order: times are identical
    bit for bit, for both update orders, with and without the refined source
    grid, for grids of one, two and four strips of 64 columns."""
    rng = np.random.default_rng(1000 + n + blk)
    h = 0.5
    L = (n - 1) * h
    if blk:
        ii = np.arange(n + 2)
        velv = np.where(((ii[:, None] // blk) + (ii[None, :] // blk)) % 2 == 0, 2.4, 3.9)
    else:
        velv = rng.uniform(2.6, 3.6, (n + 2, n + 2))
    ns, nr = 5, 9
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, L - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, L, nr)
    monkeypatch.setenv("MCT_FM2D_ORDER1", "check")   # fom 1 sweeps too
    for fom in (0, 1):
        for asgr in (0, 1):
            say("fm2d strip vs diagonal", n, blk, "fom", fom, "asgr", asgr)
            monkeypatch.setenv("MCT_SWEEP_KIND", "diag")
            want = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom, asgr=asgr, method=0)
            monkeypatch.setenv("MCT_SWEEP_KIND", "strip")
            got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, fom=fom, asgr=asgr, method=0)
            np.testing.assert_array_equal(got, want)


def test_heap_order_check_is_what_makes_order_one_exact(monkeypatch):
    """Node-scale random media with B-spline dicing (the cases above): with the
    check off, the mixed-order fixed point leaves the heap loop's answer in
    some of them (the negative control); with it on, the problems where it
    would are handed to the sequential solver and every time equals the
    oracle's."""
    monkeypatch.setenv("MCT_FM2D_ORDER1", "check")
    worst_off = worst_on = 0.0
    handed = total = 0
    for gd, sgdl, sgs in [((1, 1), 4, 8), ((2, 2), 4, 8), ((2, 1), 2, 3), ((1, 3), 5, 2)]:
        rng = np.random.default_rng(100 + 7 * gd[0] + gd[1] + sgdl + sgs)
        nvx, nvz, h = int(rng.integers(10, 30)), int(rng.integers(10, 30)), 0.6
        npd, ns, nr = 5, 6, 9
        velv = rng.uniform(2.0, 4.5, (npd, nvx + 2, nvz + 2))
        L, W = (nvx - 1) * h, (nvz - 1) * h
        ox, oz = -3.0, 1.5
        sx, sz = ox + rng.uniform(h, L - h, ns), oz + rng.uniform(h, W - h, ns)
        rx, rz = ox + rng.uniform(0, L, nr), oz + rng.uniform(0, W, nr)
        kw = dict(gdx=gd[0], gdz=gd[1], asgr=1, sgdl=sgdl, sgs=sgs, fom=1)
        want = np.stack([po.fm2d(velv[p], ox, oz, h, h, sx, sz, rx, rz, **kw) for p in range(npd)])
        monkeypatch.setenv("MCT_FM2D_VERIFY", "0")
        off = mct.modrays(velv, ox, oz, h, h, sx, sz, rx, rz, method=0, **kw)
        monkeypatch.setenv("MCT_FM2D_VERIFY", "1")
        on = mct.modrays(velv, ox, oz, h, h, sx, sz, rx, rz, method=0, **kw)
        handed += mct.modrays_last_handover()
        total += npd * ns
        worst_off = max(worst_off, float((np.abs(off - want) / want).max()))
        worst_on = max(worst_on, float((np.abs(on - want) / want).max()))
    say("check off: max rel %.3e; check on: max rel %.3e, %d of %d problems handed over" %
        (worst_off, worst_on, handed, total))
    assert worst_off > 1e-9          # the fixed point alone is not the heap loop's answer here
    assert handed > 0
    assert worst_on <= 1e-9


def test_order_one_default_path_equals_the_checked_sweeps(monkeypatch):
    """The default mixed-order path (the heap loop, one wave per problem, for
    every problem) and the sweeps with the heap-order check give the same
    times, both the oracle's, on node-scale random media with a refined source
    grid; a lane-only run (MCT_FM2D_WAVE=0) agrees bit for bit."""
    rng = np.random.default_rng(77)
    nvx, nvz, h = 41, 33, 0.5
    npd, ns, nr = 4, 8, 11
    velv = rng.uniform(2.0, 4.5, (npd, nvx + 2, nvz + 2))
    L, W = (nvx - 1) * h, (nvz - 1) * h
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, W, nr)
    kw = dict(gdx=2, gdz=2, asgr=1, sgdl=4, sgs=8, fom=1)
    want = np.stack([po.fm2d(velv[p], 0, 0, h, h, sx, sz, rx, rz, **kw) for p in range(npd)])
    default = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    monkeypatch.setenv("MCT_FM2D_WAVE", "0")
    lanes = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    monkeypatch.delenv("MCT_FM2D_WAVE")
    monkeypatch.setenv("MCT_FM2D_ORDER1", "check")
    checked = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    say("order 1 default vs checked sweeps: identical", np.array_equal(default, checked))
    np.testing.assert_array_equal(default, lanes)
    agree(default, want, 0)
    agree(checked, want, 0)
    np.testing.assert_allclose(default, checked, rtol=1e-9, atol=0)


def test_order_one_bands_wider_than_the_small_heap(monkeypatch):
    """A 64-entry heap per problem (MCT_WAVE_HEAP=64) overflows on most
    problems: the lane solver solves them again, and the times stay the
    oracle's (and equal to the default heap's)."""
    rng = np.random.default_rng(91)
    nvx, nvz, h = 61, 45, 0.5
    ns, nr = 6, 10
    velv = rng.uniform(2.2, 4.2, (nvx + 2, nvz + 2))
    L, W = (nvx - 1) * h, (nvz - 1) * h
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, W, nr)
    kw = dict(gdx=2, gdz=2, asgr=1, sgdl=4, sgs=8, fom=1)
    want = po.fm2d(velv, 0, 0, h, h, sx, sz, rx, rz, **kw)
    default = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    monkeypatch.setenv("MCT_WAVE_HEAP", "64")
    small = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    np.testing.assert_array_equal(small, default)
    agree(small, want, 0)


def test_order_one_grid_too_large_for_16_bit_heap_entries():
    """A 261 x 261 coarse grid (68,121 nodes) takes the heap-loop kernel with
    32-bit heap entries; its times are the oracle's."""
    rng = np.random.default_rng(261)
    n, h = 261, 0.4
    L = (n - 1) * h
    velv = rng.uniform(2.4, 4.0, (n + 2, n + 2))
    ns, nr = 3, 8
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, L - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, L, nr)
    for asgr in (0, 1):
        say("fm2d order 1 at 261^2, asgr", asgr)
        got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, fom=1, asgr=asgr, method=0)
        want = po.fm2d(velv, 0, 0, h, h, sx, sz, rx, rz, fom=1, asgr=asgr)
        agree(got, want, 0)


@pytest.mark.parametrize("split", ["0", "1000000"])
def test_order_one_problems_per_wave(split, monkeypatch):
    """The heap-loop kernel runs four problems per wave on large launches and
    one per wave on small ones (MCT_WAVE_SPLIT moves the boundary): both give
    the oracle's times, the same bit for bit."""
    rng = np.random.default_rng(5150)
    nvx, nvz, h = 37, 29, 0.5
    npd, ns, nr = 3, 7, 9
    velv = rng.uniform(2.2, 4.4, (npd, nvx + 2, nvz + 2))
    L, W = (nvx - 1) * h, (nvz - 1) * h
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, W, nr)
    kw = dict(gdx=2, gdz=2, asgr=1, sgdl=4, sgs=8, fom=1)
    want = np.stack([po.fm2d(velv[p], 0, 0, h, h, sx, sz, rx, rz, **kw) for p in range(npd)])
    monkeypatch.setenv("MCT_WAVE_SPLIT", split)
    got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    monkeypatch.setenv("MCT_FM2D_WAVE", "0")
    lanes = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    np.testing.assert_array_equal(got, lanes)
    agree(got, want, 0)


@pytest.mark.parametrize("split", ["0", "1000000"])
def test_order_one_heap_larger_than_lds_is_cut_to_fit(split, monkeypatch):
    """A heap of 4,096 entries asks four problems per wave for more LDS than a
    workgroup has (164 KB of 160 KB with 16-bit entries): the launcher cuts it
    to what fits, and the times are the oracle's, as with the default heap."""
    rng = np.random.default_rng(7272)
    nvx, nvz, h = 41, 33, 0.5
    npd, ns, nr = 2, 6, 9
    velv = rng.uniform(2.2, 4.4, (npd, nvx + 2, nvz + 2))
    L, W = (nvx - 1) * h, (nvz - 1) * h
    sx, sz = rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)
    rx, rz = rng.uniform(0, L, nr), rng.uniform(0, W, nr)
    kw = dict(gdx=2, gdz=2, asgr=1, sgdl=4, sgs=8, fom=1)
    want = np.stack([po.fm2d(velv[p], 0, 0, h, h, sx, sz, rx, rz, **kw) for p in range(npd)])
    say("order 1 heap 4096, split", split)
    monkeypatch.setenv("MCT_WAVE_SPLIT", split)
    monkeypatch.setenv("MCT_WAVE_HEAP", "4096")
    got = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    monkeypatch.delenv("MCT_WAVE_HEAP")
    dflt = mct.modrays(velv, 0, 0, h, h, sx, sz, rx, rz, method=0, **kw)
    np.testing.assert_array_equal(got, dflt)
    agree(got, want, 0)
