"""Served form: continuous batching of surface-wave likelihood requests.

Each result must be the same alone (a plain call) and in the mix (admitted
into batches with other chains' requests while earlier batches run), bit for
bit: every chain of a batch is computed independently.
"""
import sys
import threading
import time

import numpy as np
import pytest

import mctomo_amd as mct
from mctomo_amd import workloads as W
from mctomo_amd.workloads import _surface_sites

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def models(w, n, seed, cmin=10, cmax=60):
    rng = np.random.default_rng(seed)
    g = w.grid
    L, Z = g.dx * (g.nx - 1), g.dz * (g.nz - 1)
    out = []
    while len(out) < n:
        k = int(rng.integers(cmin, cmax + 1))
        p, q = _surface_sites(rng, k, L, Z, (2.5, 4.5))
        out.append((p, q))
    return out


def test_results_alone_equal_results_in_the_mix():
    w = W.surface_workload(n=33, nz=17, L=16.0, Z=8.0, ncells=30, np_=6, nsrc=8, nrev=8, seed=31)
    ms = models(w, 40, 7)
    plain = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=1, max_cells=60)
    alone = [plain.likelihood(p, q) for p, q in ms]
    plain.close()
    say("alone done", sum(r[0] == 0 for r in alone), "accepted of", len(alone))
    srv = mct.SurfServer(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, max_batch=8, nlanes=3,
                         order=1, max_cells=60)
    # requests from four submitting threads with random gaps: batches form from whatever is pending
    tickets = [None] * len(ms)
    rng = np.random.default_rng(3)
    gaps = rng.exponential(0.002, len(ms))

    def client(idx):
        for i in idx:
            time.sleep(gaps[i])
            tickets[i] = srv.submit(*ms[i])

    th = [threading.Thread(target=client, args=(list(range(k, len(ms), 4)),)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mixed = [srv.wait(t) for t in tickets]
    b, e, rc = srv.stats()
    srv.close()
    say("served: batches", b, "evals", e, "rc", rc)
    assert rc == 0 and e == len(ms)
    assert b < len(ms)          # some batches held more than one request
    for i, ((st_a, out_a), (st_m, out_m, lat)) in enumerate(zip(alone, mixed)):
        assert st_a == st_m, i
        if st_a == 0:
            np.testing.assert_array_equal(out_a, out_m, err_msg="request %d" % i)
        assert lat > 0


def _two_models():
    wa = W.surface_workload(n=33, nz=17, L=16.0, Z=8.0, ncells=30, np_=6, nsrc=8, nrev=8, seed=31)
    wb = W.surface_workload(n=25, nz=13, L=12.0, Z=6.0, ncells=20, np_=4, nsrc=6, nrev=6, seed=32)
    return wa, wb


def _spec(w, max_batch):
    return dict(grid=w.grid, freqs=w.freqs, src=w.src, rev=w.rev, ttime=w.ttime, raystat=w.raystat,
                order=1, max_cells=60, max_batch=max_batch)


def _alone(w, ms):
    plain = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, order=1, max_cells=60)
    out = [plain.likelihood(p, q) for p, q in ms]
    plain.close()
    return out


def test_mixed_models_alone_equal_mix():
    """Two data sets on different grids (33x33x17 with 6 periods, 25x25x13 with
    4) served by one server: requests of both arrive interleaved from several
    threads; each result equals its model's plain call bit for bit."""
    wa, wb = _two_models()
    ma, mb = models(wa, 24, 11), models(wb, 24, 12)
    alone = {0: _alone(wa, ma), 1: _alone(wb, mb)}
    srv = mct.SurfServer.mix([_spec(wa, 6), _spec(wb, 6)], max_lanes=2)
    jobs = [(0, i) for i in range(len(ma))] + [(1, i) for i in range(len(mb))]
    np.random.default_rng(5).shuffle(jobs)
    tickets = {}

    def client(part):
        for m, i in part:
            p, q = (ma if m == 0 else mb)[i]
            tickets[(m, i)] = srv.submit(p, q, model=m)

    th = [threading.Thread(target=client, args=(jobs[k::3],)) for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    res = {k: srv.wait(t) for k, t in tickets.items()}
    b, e, rc = srv.stats()
    used, peak, lanes, rel = srv.memory()
    srv.close()
    say("mix: batches", b, "evals", e, "lanes", lanes, "peak MB %.1f" % (peak / 2**20))
    assert rc == 0 and e == len(jobs)
    for (m, i), (st, out, lat) in res.items():
        st_a, out_a = alone[m][i]
        assert st == st_a, (m, i)
        if st == 0:
            np.testing.assert_array_equal(out, out_a, err_msg="model %d request %d" % (m, i))


def test_hbm_budget_moves_lanes_between_models():
    """With a device budget of exactly one lane per model, a burst for one
    model takes a second lane only by releasing the other model's idle lane;
    the footprint never exceeds the budget and every result is still the plain
    call's."""
    wa, wb = _two_models()
    probe = mct.SurfServer.mix([_spec(wa, 4), _spec(wb, 4)], max_lanes=2)
    budget, _, lanes0, _ = probe.memory()
    probe.close()
    assert lanes0 == 2 and budget > 0
    ma, mb = models(wa, 20, 21), models(wb, 20, 22)
    alone = {0: _alone(wa, ma), 1: _alone(wb, mb)}
    srv = mct.SurfServer.mix([_spec(wa, 4), _spec(wb, 4)], max_lanes=2, hbm_budget=budget)
    # a burst for model 0 (more than one lane's batch), then one for model 1
    ta = [srv.submit(p, q, model=0) for p, q in ma]
    ra = [srv.wait(t) for t in ta]
    tb = [srv.submit(p, q, model=1) for p, q in mb]
    rb = [srv.wait(t) for t in tb]
    used, peak, lanes, rel = srv.memory()
    b, e, rc = srv.stats()
    srv.close()
    say("budget %.1f MB: peak %.1f MB, lanes %d, releases %d, batches %d" % (budget / 2**20, peak / 2**20, lanes, rel, b))
    assert rc == 0 and e == 40
    assert peak <= budget
    for m, rs in ((0, ra), (1, rb)):
        for i, (st, out, lat) in enumerate(rs):
            st_a, out_a = alone[m][i]
            assert st == st_a
            if st == 0:
                np.testing.assert_array_equal(out, out_a)


def test_mix_rejects_unknown_models_and_oversized_requests():
    wa, wb = _two_models()
    srv = mct.SurfServer.mix([_spec(wa, 4), _spec(wb, 4)], max_lanes=1)
    p, q = models(wa, 1, 3)[0]
    with pytest.raises(mct.MCTError):
        srv.submit(p, q, model=2)
    big = np.zeros((61, 3))
    with pytest.raises(mct.MCTError):
        srv.submit(big, big, model=0)
    st, out, lat = srv.wait(srv.submit(p, q, model=0))
    assert st == 0
    srv.close()
    # a budget below one lane per model cannot start
    probe = mct.SurfServer.mix([_spec(wa, 4)], max_lanes=1)
    one, _, _, _ = probe.memory()
    probe.close()
    with pytest.raises(mct.MCTError):
        mct.SurfServer.mix([_spec(wa, 4), _spec(wb, 4)], max_lanes=1, hbm_budget=one)
