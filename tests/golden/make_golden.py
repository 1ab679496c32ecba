"""Write golden vectors from the reference's own code compiled in place
(oracle/_ref/libmctref.so, built by `make -C oracle ref` from /root/reference).

Only data is written: seeded inputs and the reference's outputs.  Run from the
repo root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import pyoracle as po  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def kdtree_golden():
    """kdtree2 1-NN (src/kdtree2.f90) on random and lattice (tie-heavy) sites."""
    rng = np.random.default_rng(20261202)
    sets = []
    for t in range(12):
        S = int(rng.integers(2, 420))
        pts = rng.uniform(0, 20, (S, 3))
        q = rng.uniform(-1, 21, (256, 3))
        sets.append((pts, q))
    for t in range(12):
        g = rng.integers(0, 6, (int(rng.integers(14, 160)), 3)).astype(float) * 2.0
        g = np.unique(g, axis=0)
        rng.shuffle(g)
        q = rng.integers(0, 11, (256, 3)).astype(float)
        sets.append((g, q))
    arrays = {}
    for i, (pts, q) in enumerate(sets):
        arrays["pts_%d" % i] = pts
        arrays["q_%d" % i] = q
        arrays["idx_%d" % i] = po.ref_kdtree_nearest(pts, q)
    np.savez_compressed(os.path.join(OUT, "kdtree2_nearest.npz"), n=len(sets), **arrays)


def surfdisp_golden():
    """surfdisp96 (surfmodes/surfdisp96.f) phase/group, Rayleigh/Love, on
    seeded layer models: normal dispersion, low-velocity layers, failures."""
    rng = np.random.default_rng(20261204)
    arrays = {}
    n = 0
    for trial in range(40):
        L = int(rng.integers(2, 22))
        thk = rng.uniform(0.3, 3.0, L).astype(np.float32)
        thk[-1] = 0
        vs = rng.uniform(2.0, 4.6, L)
        if trial % 2 == 0:
            vs = np.sort(vs)
        vs = vs.astype(np.float32)
        vp = (vs * 1.73).astype(np.float32)
        rho = (1.74 * vp.astype(np.float64) ** 0.25).astype(np.float32)
        per = np.geomspace(1, 5, 10) if trial % 4 < 2 else np.geomspace(0.5, 10, 20)
        for iw in (1, 2):
            for igr in (0, 1):
                cp, cg, ierr, _ = po.surfdisp96(thk, vp, vs, rho, per, iw, igr, ref=True)
                arrays["m_%d" % n] = np.stack([thk, vp, vs, rho])
                arrays["t_%d" % n] = per
                arrays["opt_%d" % n] = np.array([iw, igr, ierr])
                arrays["cp_%d" % n] = cp
                arrays["cg_%d" % n] = cg
                n += 1
    np.savez_compressed(os.path.join(OUT, "surfdisp96.npz"), n=n, **arrays)


def surfdisp_water_golden():
    """surfdisp96 on seeded models under a water layer (vp 1.5, vs 0, rho 1:
    the layer convert_to_layer puts on top when T_GRID%waterDepth > EPS,
    src/likelihood_surf.F90:32-33,546-551), Rayleigh and Love, phase and group."""
    rng = np.random.default_rng(20261213)
    arrays = {}
    n = 0
    for trial in range(24):
        L = int(rng.integers(2, 16))
        thk = rng.uniform(0.3, 3.0, L)
        thk[-1] = 0
        vs = rng.uniform(1.2, 4.6, L)
        if trial % 2 == 0:
            vs = np.sort(vs)
        vp = vs * 1.73
        rho = 1.74 * vp ** 0.25
        w = rng.uniform(0.05, 4.0)
        m = np.stack([np.r_[w, thk], np.r_[1.5, vp], np.r_[0.0, vs], np.r_[1.0, rho]]).astype(np.float32)
        per = np.geomspace(0.2, 10, 20)
        for iw in (1, 2):
            for igr in (0, 1):
                cp, cg, ierr, _ = po.surfdisp96(m[0], m[1], m[2], m[3], per, iw, igr, ref=True)
                arrays["m_%d" % n] = m
                arrays["t_%d" % n] = per
                arrays["opt_%d" % n] = np.array([iw, igr, ierr])
                arrays["cp_%d" % n] = cp
                arrays["cg_%d" % n] = cg
                n += 1
    np.savez_compressed(os.path.join(OUT, "surfdisp96_water.npz"), n=n, **arrays)


def fm2d_golden():
    """modrays phase times (fm2d/fm2dray_cartesian.f90) for one period on
    seeded smooth grids: first and mixed order, with and without the refined
    source grid.  Sources lie at least one cell inside the model (the
    reference's refined-grid exit test misbehaves at the model edge)."""
    rng = np.random.default_rng(20261205)
    arrays = {}
    n = 0
    for trial in range(10):
        nx, ny, h = int(rng.integers(10, 31)), int(rng.integers(10, 31)), 0.5
        v = rng.uniform(2.5, 4.0, (nx, ny))
        velv = np.zeros((nx + 2, ny + 2))
        velv[1:-1, 1:-1] = v
        velv[0], velv[-1] = velv[1], velv[-2]
        velv[:, 0], velv[:, -1] = velv[:, 1], velv[:, -2]
        L, W = (nx - 1) * h, (ny - 1) * h
        ns, nr = int(rng.integers(1, 5)), int(rng.integers(2, 9))
        src = np.c_[rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)]
        rec = np.c_[rng.uniform(0, L, nr), rng.uniform(0, W, nr)]
        srs = (rng.random(ns * nr) < 0.85).astype(np.int32)
        for fom in (0, 1):
            for asgr in (0, 1):
                tt = po.fm2d(velv, 0, 0, h, h, src[:, 0], src[:, 1], rec[:, 0], rec[:, 1],
                             srs=srs, fom=fom, asgr=asgr, ref=True)
                arrays["velv_%d" % n] = velv
                arrays["src_%d" % n] = src
                arrays["rec_%d" % n] = rec
                arrays["srs_%d" % n] = srs
                arrays["opt_%d" % n] = np.array([fom, asgr])
                arrays["tt_%d" % n] = tt
                n += 1
    np.savez_compressed(os.path.join(OUT, "fm2d_modrays.npz"), n=n, **arrays)


def fm2d_rays_golden():
    """rpaths ray points (fm2d/fm2dray_cartesian.f90:801-1420, uar = 0) for
    one period on seeded grids, first and mixed order, refined source grid."""
    rng = np.random.default_rng(20261206)
    arrays = {}
    n = 0
    for trial in range(8):
        nx, ny, h = int(rng.integers(12, 33)), int(rng.integers(12, 33)), 0.5
        v = rng.uniform(2.5, 4.0, (nx, ny))
        velv = np.zeros((nx + 2, ny + 2))
        velv[1:-1, 1:-1] = v
        velv[0], velv[-1] = velv[1], velv[-2]
        velv[:, 0], velv[:, -1] = velv[:, 1], velv[:, -2]
        L, W = (nx - 1) * h, (ny - 1) * h
        ns, nr = int(rng.integers(1, 4)), int(rng.integers(3, 9))
        src = np.c_[rng.uniform(h, L - h, ns), rng.uniform(h, W - h, ns)]
        rec = np.c_[rng.uniform(0, L - 1e-6, nr), rng.uniform(0, W - 1e-6, nr)]
        srs = (rng.random(ns * nr) < 0.85).astype(np.int32)
        for fom in (0, 1):
            R = po.fm2d_rays(velv, 0, 0, h, h, src[:, 0], src[:, 1], rec[:, 0], rec[:, 1],
                             srs=srs, fom=fom, asgr=1, ref=True)
            m = int(R["npts"].max())
            arrays["velv_%d" % n] = velv
            arrays["src_%d" % n] = src
            arrays["rec_%d" % n] = rec
            arrays["srs_%d" % n] = srs
            arrays["fom_%d" % n] = np.array([fom])
            arrays["npts_%d" % n] = R["npts"]
            arrays["pts_%d" % n] = R["pts"][:, :m, :]
            arrays["crazy_%d" % n] = np.array([R["crazy"]])
            n += 1
    np.savez_compressed(os.path.join(OUT, "fm2d_rpaths.npz"), n=n, **arrays)


# the reference's example data files, copied into tests/golden/formats/
FORMAT_SETS = [
    ("example2 surface", 2, "ssources.dat", "sreceivers.dat", "stimes.dat"),
    ("example2 body", 0, "bsources.dat", "breceivers.dat", "btimes.dat"),
    ("example1", 2, "ex1_sources.dat", "ex1_receivers.dat", "ex1_otimes.dat"),
    ("2dExample", 2, "ex2d_sources.dat", "ex2d_receivers.dat", "ex2d_otimes_noise.dat"),
]


def formats_golden():
    """read_data (src/likelihood_settings.f90:166-196, read_sources and
    read_times_3) on the reference's example data files."""
    d = os.path.join(OUT, "formats")
    arrays = {}
    for i, (_, dt, a, b, c) in enumerate(FORMAT_SETS):
        R = po.ref_read_data(dt, os.path.join(d, a), os.path.join(d, b), os.path.join(d, c))
        for k in ("src", "rev", "freqs", "ttime", "raystat"):
            arrays["%s_%d" % (k, i)] = R[k]
        arrays["nrays_%d" % i] = np.array([R["nrays"]])
    np.savez_compressed(os.path.join(OUT, "formats_read_data.npz"), n=len(FORMAT_SETS), **arrays)


if __name__ == "__main__":
    if po.ref() is None:
        sys.exit("oracle/_ref/libmctref.so is not built (make -C oracle ref)")
    # python make_golden.py [set ...]: only the named sets (default: all)
    sets = {"kdtree": kdtree_golden, "surfdisp": surfdisp_golden, "surfdisp_water": surfdisp_water_golden,
            "fm2d": fm2d_golden, "fm2d_rays": fm2d_rays_golden, "formats": formats_golden}
    for name in (sys.argv[1:] or list(sets)):
        sets[name]()
    print("wrote golden vectors to", OUT)
