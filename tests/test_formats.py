"""Data files (host code in the library, no GPU): read_sources / read_times_3 /
write_atimes / write_sources against the reference's own compiled readers
(src/likelihood_settings.f90, through read_data) on the reference's example
data files, on synthetic files that use Fortran list-directed input rules, and
through write -> read round trips."""
import os

import numpy as np
import pytest

import mctomo_amd as mct
from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
FMT = os.path.join(HERE, "golden", "formats")
GOLD = os.path.join(HERE, "golden", "formats_read_data.npz")
SETS = [
    (2, "ssources.dat", "sreceivers.dat", "stimes.dat"),
    (0, "bsources.dat", "breceivers.dat", "btimes.dat"),
    (2, "ex1_sources.dat", "ex1_receivers.dat", "ex1_otimes.dat"),
    (2, "ex2d_sources.dat", "ex2d_receivers.dat", "ex2d_otimes_noise.dat"),
]


@pytest.mark.parametrize("i", range(len(SETS)))
def test_example_data_files_match_reference_golden(i):
    g = np.load(GOLD)
    _, a, b, c = SETS[i]
    src, rev, fr, tt, rs, nrays = mct.read_data(os.path.join(FMT, a), os.path.join(FMT, b), os.path.join(FMT, c))
    np.testing.assert_array_equal(src, g["src_%d" % i])
    np.testing.assert_array_equal(rev, g["rev_%d" % i])
    np.testing.assert_array_equal(fr, g["freqs_%d" % i])
    np.testing.assert_array_equal(tt, g["ttime_%d" % i])
    np.testing.assert_array_equal(rs, g["raystat_%d" % i])
    assert nrays == int(g["nrays_%d" % i][0])
    # the invariants the readers establish: used pairs carry their pair index
    used = rs[:, 0, :] == 1
    idx = np.broadcast_to(np.arange(1, rs.shape[2] + 1), used.shape)
    assert np.array_equal(rs[:, 1, :][used], idx[used]) and nrays == used.sum()


def _synthetic(tmp_path, rng, quirks):
    """A random data set written with list-directed variations."""
    ns, nr, npp = int(rng.integers(1, 6)), int(rng.integers(1, 6)), int(rng.integers(1, 7))
    src = rng.uniform(-5, 25, (ns, 2))
    rev = rng.uniform(-5, 25, (nr, 3))
    fr = rng.uniform(0.05, 2.0, npp)

    def num(v):
        s = repr(float(v))
        if quirks and rng.random() < 0.3:
            s = ("%.17e" % v).replace("e", "d" if rng.random() < 0.5 else "E")
        return s

    sep = (lambda: rng.choice([" ", ",", "  ,  ", "\t"])) if quirks else (lambda: " ")
    lines = ["%d%s2" % (ns, sep())] + [sep().join(num(v) for v in p) for p in src]
    (tmp_path / "s.dat").write_text("\n".join(lines) + "\n")
    lines = ["%d %d  extra ignored" % (nr, 3) if quirks else "%d 3" % nr] + [" ".join(num(v) for v in p) for p in rev]
    (tmp_path / "r.dat").write_text("\n".join(lines) + "\n")
    out = ["%d" % npp]
    # frequencies may span records
    k = int(rng.integers(1, npp + 1)) if quirks else npp
    out.append(sep().join(num(v) for v in fr[:k]))
    if k < npp:
        if quirks:
            out.append("")
        out.append(" ".join(num(v) for v in fr[k:]))
    for _ in range(ns * nr):
        valid = int(rng.random() < 0.8)
        out.append("%d%s" % (valid, "  trailing words" if quirks and rng.random() < 0.2 else ""))
        if valid:
            for _ in range(npp):
                t = -1.0 if rng.random() < 0.15 else rng.uniform(0.5, 30.0)
                n = rng.uniform(0.01, 0.5)
                if quirks and rng.random() < 0.2:
                    out.append(num(t))
                    out.append(num(n) + " 99.0")     # the value after the item list is discarded
                else:
                    out.append(num(t) + sep() + num(n))
    (tmp_path / "t.dat").write_text("\n".join(out) + "\n")
    return [str(tmp_path / x) for x in ("s.dat", "r.dat", "t.dat")]


@pytest.mark.parametrize("quirks", [False, True])
def test_synthetic_files_match_the_compiled_reference(tmp_path, quirks):
    if po.ref() is None:
        pytest.skip("oracle/_ref not built here")
    rng = np.random.default_rng(31 if quirks else 30)
    for trial in range(25):
        d = tmp_path / ("t%d" % trial)
        d.mkdir()
        a, b, c = _synthetic(d, rng, quirks)
        R = po.ref_read_data(2, a, b, c)
        src, rev, fr, tt, rs, nrays = mct.read_data(a, b, c)
        for k, x in (("src", src), ("rev", rev), ("freqs", fr), ("ttime", tt), ("raystat", rs)):
            np.testing.assert_array_equal(x, R[k], err_msg="%s trial %d" % (k, trial))
        assert nrays == R["nrays"]


def test_write_times_round_trip(tmp_path):
    """write_atimes output reads back through read_times_3 to the same values
    (times + origin times on used periods, sigma; unused pairs marked 0)."""
    rng = np.random.default_rng(5)
    ns, nr, npp = 5, 4, 3
    npair = ns * nr
    fr = rng.uniform(0.1, 1.0, npp)
    rs = np.zeros((npp, 2, npair), np.int32)
    used = rng.random((npp, npair)) < 0.7
    used[:, 3] = False                                  # a pair no period uses
    rs[:, 0] = used
    rs[:, 1] = np.where(used, np.arange(1, npair + 1), 0)
    pt = rng.uniform(1.0, 20.0, (npp, ns, nr))
    sg = rng.uniform(0.01, 0.3, (npp, npair))
    loc = rng.uniform(-1, 1, (ns, 4))
    path = str(tmp_path / "syn.dat")
    mct.write_times(path, fr, rs, pt, sg, locations=loc)
    fr2, tt, rs2, nrays = mct.read_times(path, ns, nr)
    np.testing.assert_array_equal(fr2, fr)
    t_exp = pt.reshape(npp, npair) + np.where(used, np.repeat(loc[:, 3], nr)[None, :], 0.0)
    anyp = used.any(axis=0)
    np.testing.assert_array_equal(tt[:, 0][:, anyp], t_exp[:, anyp])
    np.testing.assert_array_equal(tt[:, 1][:, anyp], sg[:, anyp])
    assert np.all(tt[:, 0][:, ~anyp] == -1) and np.all(tt[:, 2][:, ~anyp] == 0)
    # every written time is a used pair on re-reading (no -1 was written)
    assert nrays == int(anyp.sum()) * npp
    if po.ref() is not None:
        s_path, r_path = str(tmp_path / "s.dat"), str(tmp_path / "r.dat")
        mct.write_sources(s_path, rng.uniform(0, 1, (ns, 2)))
        mct.write_sources(r_path, rng.uniform(0, 1, (nr, 2)))
        R = po.ref_read_data(2, s_path, r_path, path)
        np.testing.assert_array_equal(R["ttime"], tt)
        np.testing.assert_array_equal(R["raystat"], rs2)
        np.testing.assert_array_equal(R["src"], mct.read_sources(s_path))


def test_read_sources_bounds_and_errors(tmp_path):
    p = str(tmp_path / "s.dat")
    src = np.array([[1.0, 2.0], [3.5, -0.25]])
    mct.write_sources(p, src)
    np.testing.assert_array_equal(mct.read_sources(p), src)
    np.testing.assert_array_equal(mct.read_sources(p, bnd=[0, 4, -1, 3]), src)
    with pytest.raises(mct.MCTError) as e:
        mct.read_sources(p, bnd=[0, 3, -1, 3])       # 3.5 > 3: out_bnd
    assert e.value.code == -6
    with pytest.raises(mct.MCTError) as e:
        mct.read_sources(str(tmp_path / "missing.dat"))
    assert e.value.code == -7


def test_read_sources_checks_the_callers_capacity(tmp_path):
    """A file with more coordinates per source than the caller sized for is an
    argument error before anything is written: src holds `cap` doubles, and bnd
    must hold one (min, max) pair per coordinate of the file."""
    import ctypes
    p = str(tmp_path / "s4.dat")
    with open(p, "w") as f:
        f.write("2 4\n1 2 3 4\n5 6 7 8\n")
    L = mct.lib()
    n, nd = ctypes.c_int(), ctypes.c_int()
    guard = np.full(8, -7.0)                     # room for 2 sources x 3 coordinates, then a guard
    rc = L.mct_read_sources(p.encode(), None, 0, ctypes.c_long(6), ctypes.byref(n), ctypes.byref(nd),
                            guard.ctypes.data_as(ctypes.c_void_p))
    assert rc == -1 and n.value == 2 and nd.value == 4
    assert (guard == -7.0).all()
    bnd = np.array([0, 10, 0, 10, 0, 10.0])      # three pairs for a four-coordinate file
    out = np.zeros(8)
    rc = L.mct_read_sources(p.encode(), bnd.ctypes.data_as(ctypes.c_void_p), 6, ctypes.c_long(8),
                            ctypes.byref(n), ctypes.byref(nd), out.ctypes.data_as(ctypes.c_void_p))
    assert rc == -1
    rc = L.mct_read_sources(p.encode(), None, 0, ctypes.c_long(8), ctypes.byref(n), ctypes.byref(nd),
                            out.ctypes.data_as(ctypes.c_void_p))
    assert rc == 0 and out.tolist() == [1, 2, 3, 4, 5, 6, 7, 8]
