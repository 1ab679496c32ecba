"""Launch-shape autotuner for a surface-wave context.

The launch shape of a context (how the dispersion stage spreads columns over
waves, which 2-D sweep kernel runs and with how many threads, how many 2-D
problems go in one launch) changes the time of a batch, never its results.
The tuner times candidate shapes stage by stage on the caller's own batch —
the dispersion shapes by the dispersion stage's event time, the 2-D shapes by
the 2-D stage's — keeps the fastest of each, and checks that every candidate's
results equal the first run's bit for bit (a shape that changes a result is an
error, not a choice).  The chosen shape stays set on the context and can be
saved to, and later read from, a JSON cache keyed by the problem's shape.
"""
import json
import os

import numpy as np

from .api import SurfLikelihood

# candidates: (name, settings) with settings as (setter, value) pairs
_DSP = [
    ("lane per column", [("mode", 1)]),
    ("queue, one wave per 64 columns", [("mode", 3), (SurfLikelihood.TUNE_DSP_WAVES, 0)]),
    ("queue, 2048 waves", [("mode", 3), (SurfLikelihood.TUNE_DSP_WAVES, 2048)]),
    ("queue, 4096 waves", [("mode", 3), (SurfLikelihood.TUNE_DSP_WAVES, 4096)]),
]
_SWEEP = [
    ("strip-pipelined", [(SurfLikelihood.TUNE_SWEEP_KIND, 1)]),
    ("anti-diagonal, 64 threads", [(SurfLikelihood.TUNE_SWEEP_KIND, 2), (SurfLikelihood.TUNE_SWEEP_THREADS, 64)]),
    ("anti-diagonal, 256 threads", [(SurfLikelihood.TUNE_SWEEP_KIND, 2), (SurfLikelihood.TUNE_SWEEP_THREADS, 256)]),
]


def _apply(like, settings):
    for key, value in settings:
        if key == "mode":
            like.set_dispersion_mode(value)
        else:
            like.set_tuning(key, value)


def cache_key(like, nchain):
    g = like.grid
    return "%dx%dx%d np%d src%d rev%d chains%d" % (g.nx, g.ny, g.nz, like.np, like.nsrc, like.nrev, nchain)


def select(times):
    """The fastest candidate of a {name: [ms, ...]} table by its median time."""
    return min(times, key=lambda k: (float(np.median(times[k])), k))


def tune(like, run, outputs, nchain, repeats=2, cache=None, log=None, chunks=True):
    """Tune `like` on the caller's batch of nchain chains.  run() launches the batch and waits;
    outputs() returns its results (arrays compared bit for bit).  chunks=False
    keeps one 2-D launch per batch.  Returns a
    report {"dispersion": name, "sweep": name, "chunk": n, "times_ms": {...}}.
    With cache (a JSON path), a stored choice for this shape is applied without
    timing, and a new choice is stored."""
    key = cache_key(like, nchain)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            stored = json.load(f)
        if key in stored:
            rep = stored[key]
            _apply(like, dict(_DSP)[rep["dispersion"]])
            _apply(like, dict(_SWEEP)[rep["sweep"]])
            like.set_tuning(SurfLikelihood.TUNE_FM2D_CHUNK, rep["chunk"])
            rep = dict(rep, cached=True)
            return rep
    run()
    ref = [np.array(a, copy=True) for a in outputs()]

    def timed(stage):
        ts = []
        for _ in range(repeats):
            run()
            got = outputs()
            for a, b in zip(ref, got):
                if not np.array_equal(a, b):
                    raise RuntimeError("a launch shape changed the results")
            ts.append(float(like.timings()[stage]))
        return ts

    times = {}
    dsp = {}
    for name, settings in _DSP:
        _apply(like, settings)
        dsp[name] = timed(1)
    best_dsp = select(dsp)
    _apply(like, dict(_DSP)[best_dsp])
    times["dispersion"] = dsp
    sw = {}
    full = like.get_tuning(SurfLikelihood.TUNE_FM2D_CHUNK)
    for name, settings in _SWEEP:
        _apply(like, settings)
        sw[name] = timed(2)
    best_sw = select(sw)
    _apply(like, dict(_SWEEP)[best_sw])
    times["sweep"] = sw
    # fewer 2-D problems per launch (two launches) against all at once
    ch = {}
    for n in ((0, max(64, (full // 2 // 64) * 64)) if chunks else (0,)):
        like.set_tuning(SurfLikelihood.TUNE_FM2D_CHUNK, n)
        ch[str(n)] = timed(2)
    best_ch = int(select(ch))
    like.set_tuning(SurfLikelihood.TUNE_FM2D_CHUNK, best_ch)
    times["chunk"] = ch
    rep = {"dispersion": best_dsp, "sweep": best_sw, "chunk": best_ch, "times_ms": times, "cached": False}
    if log:
        log("autotune: dispersion %s, sweep %s, chunk %d" % (best_dsp, best_sw, best_ch))
    if cache:
        stored = {}
        if os.path.exists(cache):
            with open(cache) as f:
                stored = json.load(f)
        stored[key] = {k: rep[k] for k in ("dispersion", "sweep", "chunk")}
        with open(cache, "w") as f:
            json.dump(stored, f, indent=1)
    return rep
