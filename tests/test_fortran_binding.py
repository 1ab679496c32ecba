"""The Fortran interface module (include/mctomo_hip.f90) compiles with the
image's Fortran compiler and calls into libmctomo_hip.so through
iso_c_binding: the version string and an argument check that returns before
any device call (no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROG = """
program t
    use iso_c_binding
    use mctomo_hip
    implicit none
    type(mct_grid) :: g
    type(mct_surf_setup) :: s
    type(mct_body_setup) :: b
    type(c_ptr) :: ctx, v
    character(kind=c_char), pointer :: str(:)
    integer :: rc, n
    v = mct_version()
    call c_f_pointer(v, str, [64])
    n = 0
    do while (str(n + 1) /= c_null_char .and. n < 64)
        n = n + 1
    end do
    print '(64a)', str(1:n)
    g%nx = 0; g%ny = 0; g%nz = 0
    s%np = 0
    rc = mct_surf_create(g, s, ctx)
    print '(a,i0)', 'surf=', rc
    b%nsrc = 0
    rc = mct_body_create(g, b, ctx)
    print '(a,i0)', 'body=', rc
end program
"""


@pytest.mark.skipif(FC is None, reason="no Fortran compiler in this image")
def test_fortran_module_calls_the_library(tmp_path):
    lib = os.path.join(ROOT, "mctomo_amd")
    if not os.path.exists(os.path.join(lib, "libmctomo_hip.so")):
        pytest.skip("library not built")
    (tmp_path / "prog.f90").write_text(PROG)
    subprocess.run([FC, "-c", os.path.join(ROOT, "include", "mctomo_hip.f90"), "-o", "m.o"],
                   cwd=tmp_path, check=True, capture_output=True)
    subprocess.run([FC, "prog.f90", "m.o", "-L" + lib, "-lmctomo_hip", "-Wl,-rpath," + lib, "-o", "prog"],
                   cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./prog"], cwd=tmp_path, check=True, capture_output=True, text=True).stdout
    assert "mctomo-hip" in out
    assert "surf=-1" in out and "body=-1" in out
