"""The dispersion kernels' exp and sincos (mctomo_amd/csrc/libm_glibc.h)
return the host C library's numbers: the header is compiled for the host and
compared with glibc bit for bit on random arguments over each branch.  The
reference's surfdisp96 (and the oracle) take these values from glibc; near a
root of the period equation a last-bit difference can change surfdisp96's
root search path, so the gate is equality, not a tolerance."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not installed")
def test_exp_and_sincos_equal_glibc(tmp_path):
    exe = str(tmp_path / "libm_replica")
    src = os.path.join(HERE, "native", "libm_replica.cpp")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-mno-fma", src, "-o", exe, "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "1000000"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "total mismatches 0" in r.stdout


def test_planted_coefficient_change_is_seen(tmp_path):
    """The textbook Taylor coefficient 1/120 in place of the small-argument
    sine's minimax s2 (70 ulps apart) must show as mismatches: the comparison
    is sharp enough to tell the two apart."""
    if shutil.which("g++") is None:
        pytest.skip("g++ not installed")
    hdr = open(os.path.join(HERE, "..", "mctomo_amd", "csrc", "libm_glibc.h")).read()
    assert "s2 = 8.3333333333323288e-03" in hdr
    d = tmp_path / "csrc"
    d.mkdir()
    (d / "libm_glibc.h").write_text(hdr.replace("s2 = 8.3333333333323288e-03", "s2 = 0x1.1111111111111p-7"))
    shutil.copy(os.path.join(HERE, "..", "mctomo_amd", "csrc", "libm_glibc_tables.h"), d / "libm_glibc_tables.h")
    src = open(os.path.join(HERE, "native", "libm_replica.cpp")).read().replace(
        '"../../mctomo_amd/csrc/libm_glibc.h"', '"%s"' % str(d / "libm_glibc.h"))
    (tmp_path / "t.cpp").write_text(src)
    exe = str(tmp_path / "t")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-mno-fma", str(tmp_path / "t.cpp"),
                        "-o", exe, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "1000000"], capture_output=True, text=True, timeout=600)
    assert r.returncode != 0 and "total mismatches 0" not in r.stdout
