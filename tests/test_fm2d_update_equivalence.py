"""The parallel 2-D path's first-order update computes each square root and
quadratic solve once per node (ff_cands0f); the reference-shaped update
(ff_cands<0>, fouds1's arithmetic in the self-consistent loop) is the
specification.  Both are compiled for the host from the product source and
compared bit for bit on random stencils."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_fast_first_order_update_equals_reference_shaped_update(tmp_path):
    exe = str(tmp_path / "fuzz_fouds1")
    src = os.path.join(HERE, "native", "fuzz_fouds1.hip")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-ffp-contract=off", "-w",
           "-I" + os.path.join(HERE, "..", "mctomo_amd", "csrc"), src, "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "2000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
