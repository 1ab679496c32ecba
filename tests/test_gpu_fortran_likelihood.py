"""The m_likelihood drop-in (mctomo_amd/fortran/m_likelihood.f90) called from a
Fortran program the way the sampler calls it (src/mcmc_loc2.f90): like_setup,
likelihood, then noise_likelihood on a copy with new noise, and for body data
location_likelihood with moved sources.  The module and a mirror of the
reference's types (mctomo_amd/fortran/reference_types_mirror.f90) are compiled
with the image's Fortran compiler and linked to libmctomo_hip.so; every
printed value is compared with the oracle at the north-star gate (1e-6
relative on log-likelihoods)."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import mctomo_amd as mct
from mctomo_amd.workloads import c1_surface, body_workload, joint_workload
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(tmp):
    src = [os.path.join(ROOT, "include", "mctomo_hip.f90"),
           os.path.join(ROOT, "mctomo_amd", "fortran", "reference_types_mirror.f90"),
           os.path.join(ROOT, "mctomo_amd", "fortran", "m_likelihood.f90"),
           os.path.join(ROOT, "tests", "native", "mlike_driver.f90")]
    objs = []
    for i, f in enumerate(src):
        o = str(tmp / ("f%d.o" % i))
        r = subprocess.run([FC, "-O2", "-c", f, "-o", o, "-J" + str(tmp)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        objs.append(o)
    lib = os.path.join(ROOT, "mctomo_amd")
    exe = str(tmp / "mlike")
    r = subprocess.run([FC] + objs + ["-L" + lib, "-lmctomo_hip", "-Wl,-rpath," + lib, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


def fmt(a):
    return " ".join(repr(float(x)) for x in np.asarray(a, dtype=np.float64).reshape(-1, order="C"))


def ifmt(a):
    return " ".join(str(int(x)) for x in np.asarray(a).reshape(-1))


def write_input(path, datatype, grid, order, sigdep, points, params, dats, locations, bnoise, snoise):
    """dats: list of (datatype, freqs (np), src (nsrc, ndim), rev (nrev, 3), ttime (np,3,npair),
    raystat (np,2,npair)) in the library's C layouts (= the reference's Fortran arrays)."""
    g = grid
    L = ["%d %d" % (datatype, len(dats)),
         "%d %d %d %r %r %r %r %r %r %r %r %r 0.0 %r" % (
             g.nx, g.ny, g.nz, g.xmin, g.xmin + g.dx * (g.nx - 1), g.ymin, g.ymin + g.dy * (g.ny - 1),
             g.zmin, g.zmin + g.dz * (g.nz - 1), g.dx, g.dy, g.dz, g.scaling),
         "%d 0 1 0 1 1 1 4 8 %d 0" % (sigdep, order),
         "1.0e-3 0.5",
         "%d" % points.shape[0], fmt(points), fmt(params)]
    for dt, fr, src, rev, tt, rs in dats:
        npair = src.shape[0] * rev.shape[0]
        nrays = int((rs[:, 0, :] == 1).sum())
        L += ["%d %d %d %d %d %d" % (dt, len(fr), src.shape[0], rev.shape[0], nrays, src.shape[1]),
              fmt(src), fmt(rev), fmt(fr), fmt(tt), ifmt(rs)]
        assert tt.shape[2] == npair
    L += ["%d" % (0 if locations is None else locations.shape[0])]
    if locations is not None:
        L.append(fmt(locations))
    L += ["%d" % (0 if bnoise is None else bnoise.shape[1])]
    if bnoise is not None:
        L.append(fmt(bnoise[0]) + " " + fmt(bnoise[1]))
    L += ["%d" % (0 if snoise is None else snoise.shape[1])]
    if snoise is not None:
        L.append(fmt(snoise[0]) + " " + fmt(snoise[1]))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def run(exe, path):
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    res = {}
    for line in r.stdout.splitlines():
        f = line.split()
        if f and f[0] in ("LIKE", "NOISE", "LOCATION"):
            res[f[0]] = np.array([float(x) for x in f[1:4]])
        elif f and f[0] == "PART":
            res["PART%s" % f[1]] = np.array([float(x) for x in f[2:5]])
        elif f and f[0] == "BAD":
            res["BAD"] = (int(f[1]), int(f[2]))
    say(r.stdout)
    return res


def gdict(g):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin, dx=g.dx, dy=g.dy, dz=g.dz,
                scaling=g.scaling)


def close(got, want, tol=1e-6):
    assert np.all(np.abs(got - want) <= tol * np.abs(want)), (got, want)


def bumped(n):
    return n * 1.5 + 0.01


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if FC is None:
        pytest.skip("no Fortran compiler in this image")
    return build(tmp_path_factory.mktemp("mlike"))


def test_surface_c1_through_the_fortran_module(exe, tmp_path):
    w = c1_surface(order=1)
    npp = w.freqs.shape[0]
    sn = np.stack([np.full(npp, 0.01), np.full(npp, 0.04)])
    path = str(tmp_path / "c1.in")
    src3 = np.c_[w.src, np.zeros(w.nsrc)]
    rev3 = np.c_[w.rev, np.zeros(w.rev.shape[0])]
    write_input(path, 2, w.grid, 1, 1, w.points, w.params, [(2, w.freqs, src3, rev3, w.ttime, w.raystat)],
                None, None, sn)
    res = run(exe, path)
    grid = gdict(w.grid)
    o1 = po.surf_likelihood(grid, w.points, w.params, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=1,
                            sigdep=1, snoise0=sn[0], snoise1=sn[1])
    o2 = po.surf_likelihood(grid, w.points, w.params, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=1,
                            sigdep=1, snoise0=bumped(sn[0]), snoise1=bumped(sn[1]))
    assert o1[0] == 0 and o2[0] == 0
    close(res["LIKE"], o1[1])
    close(res["NOISE"], o2[1])
    assert res["BAD"] == (0, 0)


def test_body_p_and_s_through_the_fortran_module(exe, tmp_path):
    w = body_workload(n=29, h=0.3, ncells=30, nsrc=5, nrev=12, datatype=1, seed=5)
    bn = np.array([[0.01, 0.02], [0.05, 0.06]])
    path = str(tmp_path / "b.in")
    src4 = w.locations.copy()
    write_input(path, 1, w.grid, 0, 1, w.points, w.params, [(1, np.zeros(2), src4, w.rev, w.ttime, w.raystat)],
                w.locations, bn, None)
    res = run(exe, path)
    grid = gdict(w.grid)
    _, vp, vs, _ = po.kdtree_to_grid(grid, w.points, w.params)
    pt = np.stack([po.body_traveltimes(v, grid, w.locations, w.rev, 0) for v in (vp, vs)])
    o1, *_ = po.body_misfit(pt, w.ttime, w.raystat, w.locations, sigdep=1, bnoise0=bn[0], bnoise1=bn[1])
    o2, *_ = po.body_misfit(pt, w.ttime, w.raystat, w.locations, sigdep=1, bnoise0=bumped(bn[0]),
                            bnoise1=bumped(bn[1]))
    loc2 = w.locations.copy()
    loc2[:, 0:3] += 0.05
    loc2[:, 3] += 0.02
    pt3 = np.stack([po.body_traveltimes(v, grid, loc2, w.rev, 0) for v in (vp, vs)])
    o3, *_ = po.body_misfit(pt3, w.ttime, w.raystat, loc2, sigdep=1, bnoise0=bumped(bn[0]), bnoise1=bumped(bn[1]))
    close(res["LIKE"], o1)
    close(res["NOISE"], o2)
    close(res["LOCATION"], o3)


def test_joint_through_the_fortran_module(exe, tmp_path):
    w = joint_workload(n=33, nz=17, ncells=30, nsrc=5, nrev=32, np_=5, nsta=6, seed=43)
    bn = np.array([[0.01, 0.02], [0.05, 0.06]])
    npp = len(w.surf["freqs"])
    sn = np.stack([np.full(npp, 0.01), np.full(npp, 0.04)])
    b, s = w.body, w.surf
    path = str(tmp_path / "j.in")
    sta3 = np.c_[s["src"], np.zeros(len(s["src"]))]
    write_input(path, 3, w.grid, 0, 1, w.points, w.params,
                [(1, np.zeros(2), w.locations, b["rev"], b["ttime"], b["raystat"]),
                 (2, s["freqs"], sta3, sta3, s["ttime"], s["raystat"])], w.locations, bn, sn)
    res = run(exe, path)
    grid = gdict(w.grid)
    _, vp, vs, _ = po.kdtree_to_grid(grid, w.points, w.params)
    pt = np.stack([po.body_traveltimes(v, grid, w.locations, b["rev"], 0) for v in (vp, vs)])
    ob, *_ = po.body_misfit(pt, b["ttime"], b["raystat"], w.locations, sigdep=1, bnoise0=bn[0], bnoise1=bn[1])
    rc, os_, *_ = po.surf_likelihood(grid, w.points, w.params, s["freqs"], s["src"], s["rev"], s["ttime"],
                                     s["raystat"], order=0, sigdep=1, snoise0=sn[0], snoise1=sn[1], datatype=3)
    assert rc == 0
    close(res["PART1"], ob)
    close(res["PART2"], os_)
    close(res["LIKE"], ob + os_)
    ob2, *_ = po.body_misfit(pt, b["ttime"], b["raystat"], w.locations, sigdep=1, bnoise0=bumped(bn[0]),
                             bnoise1=bumped(bn[1]))
    _, os2, *_ = po.surf_likelihood(grid, w.points, w.params, s["freqs"], s["src"], s["rev"], s["ttime"],
                                    s["raystat"], order=0, sigdep=1, snoise0=bumped(sn[0]), snoise1=bumped(sn[1]),
                                    datatype=3)
    close(res["NOISE"], ob2 + os2)
