"""The launch-shape autotuner's host logic (mctomo_amd/autotune.py) against a
stand-in context whose stage times depend on the shape: it keeps the fastest
shape of each stage, leaves it set, refuses a shape that changes results, and
stores and reuses its choice through the cache."""
import numpy as np
import pytest

from mctomo_amd import autotune as AT
from mctomo_amd.api import SurfLikelihood as SL


class FakeGrid:
    nx, ny, nz = 9, 8, 7


class FakeCtx:
    """Stage times: dispersion fastest with 2048 queue waves, 2-D fastest with
    the 256-thread anti-diagonal sweep and two launches."""
    grid = FakeGrid()
    np, nsrc, nrev = 4, 3, 5

    def __init__(self, corrupt=None):
        self.mode, self.cfg, self.corrupt = 0, {1: 0, 2: 0, 3: 0, 4: 0}, corrupt
        self.full = 1280
        self.calls = 0

    def set_dispersion_mode(self, m):
        self.mode = m

    def set_tuning(self, k, v):
        self.cfg[k] = v

    def get_tuning(self, k):
        return self.full if k == SL.TUNE_FM2D_CHUNK and self.cfg[k] == 0 else self.cfg[k]

    def timings(self):
        dsp = {1: 30.0, 3: 20.0}[self.mode] - (5.0 if self.cfg[1] == 2048 else 0.0)
        sw = 50.0 if self.cfg[2] == 1 else (40.0 if self.cfg[3] == 256 else 60.0)
        if 0 < self.cfg[4] < self.full:
            sw -= 1.0
        return np.array([1.0, dsp, sw, 0.1])

    def run(self):
        self.calls += 1

    def outputs(self):
        out = np.arange(6.0)
        if self.corrupt and self.cfg[2] == self.corrupt:
            out[3] += 1e-12
        return [out]


def test_tuner_keeps_the_fastest_shape_of_each_stage(tmp_path):
    c = FakeCtx()
    rep = AT.tune(c, c.run, c.outputs, nchain=4)
    assert rep["dispersion"] == "queue, 2048 waves"
    assert rep["sweep"] == "anti-diagonal, 256 threads"
    assert rep["chunk"] == 640
    assert c.mode == 3 and c.cfg[1] == 2048 and c.cfg[2] == 2 and c.cfg[3] == 256 and c.cfg[4] == 640
    # the cache stores the choice and a new context takes it without timing
    path = str(tmp_path / "tune.json")
    AT.tune(FakeCtx(), FakeCtx().run, FakeCtx().outputs, nchain=4, cache=path)
    d = FakeCtx()
    rep2 = AT.tune(d, d.run, d.outputs, nchain=4, cache=path)
    assert rep2["cached"] and d.calls == 0
    assert d.mode == 3 and d.cfg[1] == 2048 and d.cfg[2] == 2 and d.cfg[4] == 640


def test_tuner_refuses_a_shape_that_changes_results():
    c = FakeCtx(corrupt=2)
    with pytest.raises(RuntimeError):
        AT.tune(c, c.run, c.outputs, nchain=4)


def test_select_breaks_ties_by_name():
    assert AT.select({"b": [1.0, 3.0], "a": [2.0, 2.0], "c": [5.0]}) == "a"
