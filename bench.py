#!/usr/bin/env python3
"""bench.py -- likelihood evaluations per second on one MI355X.

Default workload (--workload surf201): the surface-wave likelihood at the
201x201x61 grid of the served mix -- 100 x 100 x 30 km at 0.5 km, Voronoi
sites with Vs in [2.5,4.5] km/s (slower sites shallower so every model passes
check_model), Rayleigh fundamental-mode phase velocity per column at 20
periods log-spaced 1-10 s, then 2-D phase traveltimes per period from 32
sources to 32 receivers (refined source grid 4 x +-8), Gaussian misfit.
--order 0 (the default) is the first-order upwind update, --order 1 the
mixed-order update the reference's example settings use; the JSON line
carries an order-1 batched figure beside the order-0 `value`.  One evaluation = one likelihood() call with the whole grid as the
perturbed box: nearest site (2.46 M nodes), 40,401 dispersion curves, 640
2-D solves of 40,401 nodes, 20,480 picks.

Plain: one chain per call through the host entry point (the drop-in), timed
per call for p50/p99 latency.  Batched: B chains per launch on device arrays
(inputs already in HBM), Voronoi cell counts uniform in [20,400]; a step is
one batch.  `value` is the batched evaluations per second.

--workload body101: the 3-D body-wave eikonal drop-in (100 sites, 101^3 at
0.2 km, 50 sources x 400 receivers).
--workload c3: the batched configuration (256 chains, 101x101x41, 20 periods
0.5-10 s, 64 x 64 pairs, Rayleigh and Love, phase and group velocities; an
eval is a chain's four likelihoods).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "likelihood evals/sec, 1 GPU (batched/served vs plain) + p99 eval latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def ctypes_vp(p):
    import ctypes
    return ctypes.c_void_p(p)


def ctypes_i64(n):
    import ctypes
    return ctypes.c_int64(n)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gdict(g):
    return dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin, dx=g.dx,
                dy=g.dy, dz=g.dz, scaling=g.scaling)


# --------------------------------------------------------------------- surface
def surf_chains(w, nchain, cmin, cmax, seed):
    """Per-chain Voronoi models on the workload's data: own sites, cell counts
    uniform in [cmin, cmax], Vs ranked by site depth.  Chain 0 is the
    workload's own model."""
    from mctomo_amd.workloads import _surface_sites
    rng = np.random.default_rng(seed)
    g = w.grid
    L = g.dx * (g.nx - 1)
    Z = g.dz * (g.nz - 1)
    P = np.zeros((nchain, cmax, 3))
    Q = np.zeros((nchain, cmax, 3))
    N = np.zeros(nchain, np.int32)
    for c in range(nchain):
        if c == 0:
            p, q = w.points, w.params
        else:
            k = int(rng.integers(cmin, cmax + 1))
            p, q = _surface_sites(rng, k, L, Z, (2.5, 4.5))
        P[c, :len(p)], Q[c, :len(p)], N[c] = p, q, len(p)
    return P, Q, N, rng


def run_surface(args):
    import mctomo_amd as mct
    from mctomo_amd.workloads import surface_workload, _surface_sites

    w = surface_workload(201, 61, 100.0, 30.0, 200, 20, (1.0, 10.0), 32, 32, seed=20261206,
                         order=args.order)
    B = max(1, args.batch)
    cmin, cmax = 20, 400
    like = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=1,
                              order=args.order, max_cells=cmax, max_chains=B)
    like.set_method(args.fm2d_method)
    P, Q, N, rng = surf_chains(w, B, cmin, cmax, 4242)
    dP, dQ, dN = (mct.DeviceArray.from_numpy(a) for a in (P, Q, N))
    dS0 = mct.DeviceArray.from_numpy(np.zeros((B, 2, w.freqs.shape[0])))
    dO = mct.DeviceArray((B, 3), np.float64)
    dS = mct.DeviceArray((B,), np.int32)

    def run(nchain):
        like.likelihood_batch_dev(nchain, dP.ptr, dQ.ptr, dN.ptr, dS0.ptr, dO.ptr, dS.ptr)
        like.sync()

    # every chain's model must pass check_model and dispersion (a rejected
    # model skips the 2-D solves): redraw rejected chains, checked on the GPU
    for attempt in range(20):
        run(B)
        st = dS.numpy()
        bad = np.flatnonzero(st != 0)
        if bad.size == 0:
            break
        if (st < 0).any():
            raise RuntimeError("chain status %s" % st)
        L_ = w.grid.dx * (w.grid.nx - 1)
        Z_ = w.grid.dz * (w.grid.nz - 1)
        for c in bad:
            k = int(rng.integers(cmin, cmax + 1))
            p, q = _surface_sites(rng, k, L_, Z_, (2.5, 4.5))
            P[c] = 0; Q[c] = 0
            P[c, :k], Q[c, :k], N[c] = p, q, k
        dP.upload(P); dQ.upload(Q); dN.upload(N)
    else:
        raise RuntimeError("could not draw %d valid chains" % B)

    # ---- plain single-request drop-in ------------------------------------------
    for _ in range(args.warmup):
        like.likelihood(w.points, w.params)
    lat = []
    for _ in range(args.plain_evals):
        t0 = time.perf_counter()
        rc, out = like.likelihood(w.points, w.params)
        lat.append(time.perf_counter() - t0)
        if rc != 0:
            raise RuntimeError("plain call status %d" % rc)
    lat = np.array(lat)
    plain_eps = 1.0 / lat.mean()
    st_plain = like.timings()
    log("plain: mean %.2f ms p50 %.2f p99 %.2f  stages(ms) nearest %.2f dispersion %.2f "
        "2-D %.2f misfit %.2f" % (lat.mean() * 1e3, np.median(lat) * 1e3,
                                  np.quantile(lat, 0.99) * 1e3, *st_plain))
    ev1, lev1 = like.dispersion_work(1)
    log("2-D sweep (plain): slots %d, per problem: rounds %.2f, diagonal steps %.0f, node updates %.0f"
        % like.fm2d_stats())

    # ---- launch shape picked by the autotuner (results unchanged, checked) ----------
    tuned = None
    if not args.no_autotune:
        from mctomo_amd import autotune
        # one 2-D launch per batch, so that the roofline below times a whole launch
        rep = autotune.tune(like, lambda: run(B), lambda: [dO.numpy(), dS.numpy()], B, repeats=1, log=log,
                            chunks=False)
        tuned = {k: rep[k] for k in ("dispersion", "sweep", "chunk")}
        tuned["stage_ms"] = {st: {k: round(float(np.median(v)), 2) for k, v in tab.items()}
                             for st, tab in rep["times_ms"].items()}

    # ---- batched: the timed region -----------------------------------------------
    for _ in range(args.warmup):
        run(B)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(B)
    dt = time.perf_counter() - t0
    if (dS.numpy() != 0).any():
        raise RuntimeError("non-zero chain status %s" % dS.numpy())
    st_b = like.timings()
    evB, levB = like.dispersion_work(B)
    ms_step = dt / args.steps * 1e3
    batched_eps = B * args.steps / dt
    log("batched: B=%d %.1f ms/step %.2f evals/s  stages(ms) nearest %.2f dispersion %.2f "
        "2-D %.2f misfit %.2f" % (B, ms_step, batched_eps, *st_b))
    log("2-D stage: near-source %.2f ms, sweep %.2f ms" % like.timings2d())

    # ---- the same batch with the mixed-order update (the reference examples'
    # LIKE_SET%ORDER = 1): its 2-D times depend on the heap loop's pop order,
    # so the 2-D stage runs the heap loop itself, one wave per problem
    order1 = None
    if args.order == 0 and args.order1_steps > 0:
        like1 = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=1, order=1,
                                   max_cells=cmax, max_chains=B)
        dO1 = mct.DeviceArray((B, 3), np.float64)
        dS1 = mct.DeviceArray((B,), np.int32)

        def run1():
            like1.likelihood_batch_dev(B, dP.ptr, dQ.ptr, dN.ptr, dS0.ptr, dO1.ptr, dS1.ptr)
            like1.sync()
        run1()
        t0 = time.perf_counter()
        for _ in range(args.order1_steps):
            run1()
        dt1 = time.perf_counter() - t0
        st1 = dS1.numpy()
        checked, handed = like1.fm2d_handover()
        order1 = {"evals_per_s": round(B * args.order1_steps / dt1, 3), "steps": args.order1_steps,
                  "ms_per_step": round(dt1 / args.order1_steps * 1e3, 3),
                  "statuses": {str(k): int(v) for k, v in zip(*np.unique(st1, return_counts=True))},
                  "stage_ms": [round(v, 3) for v in like1.timings()],
                  "problems": int(B * w.freqs.shape[0] * w.nsrc),
                  "solved_by_heap_loop": int(handed), "failed_heap_order_check": int(checked)}
        log("order 1 batched:", order1)

    # ---- perturbed-box proposals (the secondary unit) ----------------------------
    # One proposal per chain: the Vs of the chain's deepest cell rises by 0.05
    # km/s; the perturbed box is that cell's Voronoi region (its nodes in the GPU
    # nearest-site grid), what the sampler's box covers for a velocity change.
    # Dispersion runs on the box's columns widened by one, the 2-D solves and the
    # misfit on the whole grid; the batch above is the accepted state.
    pbox = None
    if args.box_steps > 0:
        like.accept_batch(np.ones(B, np.int32))
        Q2 = Q.copy()
        cols = []
        for c in range(B):
            k = int(np.argmax(P[c, :N[c], 2]))
            sid, _, _, _ = mct.kdtree_to_grid(w.grid, P[c, :N[c]], Q[c, :N[c]])
            ix, iy = np.nonzero(np.any(sid == k + 1, axis=2))
            e = 1e-9
            box = [w.grid.xmin + ix.min() * w.grid.dx + e, w.grid.ymin + iy.min() * w.grid.dy + e, 0.0,
                   w.grid.xmin + ix.max() * w.grid.dx + e, w.grid.ymin + iy.max() * w.grid.dy + e, 0.0]
            cols.append(like.box_columns(box))
            Q2[c, k, 1] += 0.05
            Q2[c, k, 0] = 1.73 * Q2[c, k, 1]
        cols = np.array(cols, np.int32)
        dQ2 = mct.DeviceArray.from_numpy(Q2)

        def run_box():
            like.likelihood_box_batch_dev(B, dP.ptr, dQ2.ptr, dN.ptr, cols, dS0.ptr, dO.ptr, dS.ptr)
            like.sync()

        run_box()
        o_box, s_box = dO.numpy().copy(), dS.numpy().copy()
        st_box = like.timings()
        t0 = time.perf_counter()
        for _ in range(args.box_steps):
            run_box()
        dtb = time.perf_counter() - t0
        like.likelihood_batch_dev(B, dP.ptr, dQ2.ptr, dN.ptr, dS0.ptr, dO.ptr, dS.ptr)
        like.sync()
        same = bool(np.array_equal(o_box, dO.numpy()) and np.array_equal(s_box, dS.numpy()))
        frac = float(np.mean([(b[1] - b[0] + 1) * (b[3] - b[2] + 1) for b in cols]) / (w.grid.nx * w.grid.ny))
        pbox = {"evals_per_s": round(B * args.box_steps / dtb, 3), "steps": args.box_steps,
                "mean_box_column_fraction": round(frac, 4), "statuses_ok": bool((s_box == 0).all()),
                "equal_to_whole_grid_call": same, "stage_ms": [round(v, 3) for v in st_box],
                "proposal": "Vs of each chain's deepest cell + 0.05 km/s; box = that cell's Voronoi region"}
        log("perturbed box:", pbox)

    # ---- group velocities (phaseGroup = 1, a secondary figure) -----------------
    # The same chains with group times: surfdisp96's group pass, rays from every
    # receiver (rpaths) and the group delays along them (CalGroupTime).
    group = None
    if args.group_steps > 0:
        lg = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=1,
                                order=args.order, max_cells=cmax, max_chains=B, phase_group=1)
        lg.set_method(args.fm2d_method)

        def run_group():
            lg.likelihood_batch_dev(B, dP.ptr, dQ.ptr, dN.ptr, dS0.ptr, dO.ptr, dS.ptr)
            lg.sync()

        run_group()
        t0 = time.perf_counter()
        for _ in range(args.group_steps):
            run_group()
        dtg = time.perf_counter() - t0
        sg_ = dS.numpy()
        group = {"evals_per_s": round(B * args.group_steps / dtg, 3), "steps": args.group_steps,
                 "statuses": {str(int(v)): int((sg_ == v).sum()) for v in np.unique(sg_)},
                 "stage_ms": [round(v, 3) for v in lg.timings()]}
        log("group velocities:", group)
        lg.close()

    # ---- served: Poisson arrivals into the continuous-batching server -----------
    served = None
    if args.served > 0:
        srv = mct.SurfServer(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, max_batch=min(B, 32), nlanes=2,
                             raylov=1, order=args.order, max_cells=cmax)
        srv.set_fill(args.served_fill_ms)
        reqs = [(P[c % B, :N[c % B]], Q[c % B, :N[c % B]]) for c in range(args.served)]
        for t in [srv.submit(*reqs[i]) for i in range(min(B, len(reqs)))]:
            srv.wait(t)                                   # warm-up
        # the server's own capacity (a closed burst of B requests), then Poisson
        # arrivals at 0.8x it
        t0 = time.perf_counter()
        for t in [srv.submit(*reqs[i]) for i in range(min(B, len(reqs)))]:
            srv.wait(t)
        cap_s = min(B, len(reqs)) / (time.perf_counter() - t0)
        rate = 0.8 * cap_s
        gaps = np.random.default_rng(7).exponential(1.0 / rate, len(reqs))
        tickets, t_in = [], []
        t0 = time.perf_counter()
        for i, (p_, q_) in enumerate(reqs):
            t_next = t0 + gaps[:i + 1].sum()
            now = time.perf_counter()
            if t_next > now:
                time.sleep(t_next - now)
            t_in.append(time.perf_counter() - t0)
            tickets.append(srv.submit(p_, q_))
        t_sub = time.perf_counter() - t0
        res_s = [srv.wait(t) for t in tickets]
        t_all = time.perf_counter() - t0
        lat_s = np.array([r[2] for r in res_s])
        # requests still in the server when the last one arrived: a server that keeps
        # up holds about a batch or two; one that does not holds a backlog that grows
        # with the run
        backlog = int(sum(ti + li * 1e-3 > t_sub for ti, li in zip(t_in, lat_s)))
        bad = [r[0] for r in res_s if r[0] != 0]
        nb, ne, rc = srv.stats()
        srv.close()
        if bad or rc:
            raise RuntimeError("served statuses %s rc %d" % (bad[:5], rc))
        served = {"requests": len(reqs), "closed_burst_capacity_per_s": round(cap_s, 3),
                  "arrival_rate_per_s": round(rate, 3),
                  "offered_per_s": round(len(reqs) / t_sub, 3), "evals_per_s": round(len(reqs) / t_all, 3),
                  "backlog_at_last_arrival": backlog,
                  "p50_ms": round(float(np.median(lat_s)), 3), "p99_ms": round(float(np.quantile(lat_s, 0.99)), 3),
                  "batches": int(nb - 1), "lanes": 2, "max_batch": min(B, 32), "fill_ms": args.served_fill_ms}
        log("served:", served)

    # ---- served mix: three grids in one server, 1,024 chains -------------------
    # One request from each of 1,024 chains, each chain on one of three data
    # sets (51x51x21 with 10 periods and 16x16 pairs, 101x101x41 with 20 periods
    # and 32x32, 201x201x61 with 30 periods and 32x32), Voronoi cell counts
    # U{10..800}.  A closed burst of 192 requests measures the server's
    # capacity; then the 1,024 requests arrive as a Poisson process at 0.8x it.
    # Sustained = requests / (last result - first arrival).
    mixres = None
    if args.served_mix > 0:
        from mctomo_amd.workloads import surface_workload as sw_
        wm = [sw_(51, 21, 25.0, 10.0, 50, 10, (1.0, 5.0), 16, 16, seed=20261210, order=args.order),
              sw_(101, 41, 50.0, 20.0, 100, 20, (1.0, 10.0), 32, 32, seed=20261211, order=args.order),
              sw_(201, 61, 100.0, 30.0, 200, 30, (1.0, 10.0), 32, 32, seed=20261212, order=args.order)]
        mixc = 800
        mbs = (64, 64, 32)
        specs = [dict(grid=x.grid, freqs=x.freqs, src=x.src, rev=x.rev, ttime=x.ttime, raystat=x.raystat,
                      raylov=1, order=args.order, max_cells=mixc, max_batch=mb) for x, mb in zip(wm, mbs)]
        srvm = mct.SurfServer.mix(specs, max_lanes=2, hbm_budget=args.served_hbm_gb << 30)
        srvm.set_fill(args.served_fill_ms)
        rng_m = np.random.default_rng(99)

        def draw(m):
            g_ = wm[m].grid
            k = int(rng_m.integers(10, mixc + 1))
            return _surface_sites(rng_m, k, g_.dx * (g_.nx - 1), g_.dz * (g_.nz - 1), (2.5, 4.5))

        nreq = args.served_mix
        nburst = 192
        mreq = [int(v) for v in rng_m.integers(0, 3, nburst + nreq)]
        reqs_m = [(m,) + draw(m) for m in mreq]
        burst, live = reqs_m[:nburst], reqs_m[nburst:]
        for t in [srvm.submit(p_, q_, model=m) for m, p_, q_ in burst[:48]]:
            srvm.wait(t)                                  # lanes built, first graphs captured
        t0 = time.perf_counter()
        for t in [srvm.submit(p_, q_, model=m) for m, p_, q_ in burst]:
            srvm.wait(t)
        cap = len(burst) / (time.perf_counter() - t0)
        rate_m = 0.8 * cap
        arrive = np.cumsum(np.random.default_rng(7).exponential(1.0 / rate_m, len(live)))
        tickets, t_in = [], []
        t0 = time.perf_counter()
        for i, (m, p_, q_) in enumerate(live):
            now = time.perf_counter()
            if t0 + arrive[i] > now:
                time.sleep(t0 + arrive[i] - now)
            t_in.append(time.perf_counter() - t0)
            tickets.append((m, srvm.submit(p_, q_, model=m)))
        t_sub = time.perf_counter() - t0
        res_m = [(m, srvm.wait(t)) for m, t in tickets]
        t_all = time.perf_counter() - t0
        used_b, peak_b, lanes_m, rel_m = srvm.memory()
        nb_m, ne_m, rc_m = srvm.stats()
        srvm.close()
        if rc_m or any(r[0] < 0 for _, r in res_m):
            raise RuntimeError("served mix rc %d" % rc_m)
        lat_m = np.array([r[2] for _, r in res_m])
        backlog_m = int(sum(ti + li * 1e-3 > t_sub for ti, li in zip(t_in, lat_m)))
        per = {}
        for mm, name in enumerate(("51x51x21/10", "101x101x41/20", "201x201x61/30")):
            lm = np.array([r[2] for m, r in res_m if m == mm])
            if lm.size:
                per[name] = {"requests": int(lm.size), "p50_ms": round(float(np.median(lm)), 3),
                             "p99_ms": round(float(np.quantile(lm, 0.99)), 3)}
        sustained = len(live) / t_all
        offered = len(live) / t_sub
        mixres = {"chains": len(live), "closed_burst_capacity_per_s": round(cap, 3),
                  "arrival_rate_per_s": round(rate_m, 3), "offered_per_s": round(offered, 3),
                  "evals_per_s": round(sustained, 3), "sustained_over_offered": round(sustained / offered, 4),
                  "backlog_at_last_arrival": backlog_m,
                  "p50_ms": round(float(np.median(lat_m)), 3), "p99_ms": round(float(np.quantile(lat_m, 0.99)), 3),
                  "rejected_models": int(sum(r[0] > 0 for _, r in res_m)), "per_grid": per,
                  "lanes": lanes_m, "max_batch": list(mbs), "fill_ms": args.served_fill_ms,
                  "hbm_budget_gb": args.served_hbm_gb, "device_bytes_peak": int(peak_b),
                  "lane_releases": int(rel_m), "batches": int(nb_m)}
        log("served mix:", mixres)

    # ---- roofline of the dominant kernel, live HIP events ------------------------
    # Two candidates, timed by events on the stream they run on: the dispersion
    # launch k_dispersion_sm (FP64 VALU; algorithmic flops = 170 per dltar
    # evaluation x layer step, Rayleigh) and the 2-D sweep launch k_fm2d_strip
    # (streams the time and velocity grids; algorithmic bytes per solve =
    # 8 C velocities read + 8 C times written + 8 per receiver pick).  A batch
    # may run in parts; the stage times follow the first part.
    nparts, n0 = like.split_parts(B)
    ev0, lev0 = like.dispersion_work(n0)
    sc0, slev0 = like.dispersion_solved(n0)
    G = w.grid.nx * w.grid.ny * w.grid.nz
    C = w.grid.nx * w.grid.ny
    npd = w.freqs.shape[0]
    nsrc, nrev = w.src.shape[0], w.rev.shape[0]
    t_near, t_sweep = like.timings2d()
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_surf201.json")
    try:
        with open(prof) as f:
            traffic = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        traffic = {}

    def traffic_of(kernel, chains):
        # HBM bytes per launch from the committed counter profile of this workload
        # (FETCH_SIZE and WRITE_SIZE, scaled by a copy of known size), per chain
        kt = traffic.get(kernel)
        return int(kt["bytes_per_chain"] * chains) if kt else None

    flops = 170.0 * slev0
    a_disp = flops / (st_b[1] * 1e-3) / 1e12
    r_disp = {"bound": "fp64", "achieved": round(a_disp, 4), "peak": FP64_PEAK_TFLOPS,
              "unit": "TFLOP/s", "frac": round(a_disp / FP64_PEAK_TFLOPS, 6),
              "traffic": traffic_of("k_dispersion_sm", n0),
              "kernel": "k_dispersion_sm", "launch_chains": int(n0), "launch_ms": round(st_b[1], 3),
              "note": "FP64 vector arithmetic (no MFMA; VALU-bound): the peak quoted is MI355X's FP64 "
                      "figure (78.6 TFLOP/s, vector and matrix alike); flops = 170 x sum over the columns "
                      "solved of dltar evaluations x (layers-1) (the sequential root walk's count); "
                      "columns whose Vs profile repeats an earlier column of the chain are copied",
              "columns": int(n0 * C), "columns_solved": int(sc0),
              "dltar_evals": int(ev0), "layer_evals": int(lev0), "layer_evals_solved": int(slev0),
              "alg_bytes": int(n0 * (4 * G + 8 * C * npd))}
    solves = n0 * npd * nsrc
    # the survey's fixed definition for a 2-D solve: 4 (C + R) + 4 Nr bytes (FP32
    # times over the coarse and refined nodes, and the picks); the strip kernel
    # sweeps the coarse grid and picks, so its share is 4 C + 4 Nr
    alg_sweep = solves * (4.0 * C + 4.0 * nrev)
    r_sweep = None
    heap1 = args.order == 1 and os.environ.get("MCT_FM2D_ORDER1", "") != "check"
    if t_sweep > 0 and heap1:
        # order 1: the 2-D stage is the reference's heap loop, four problems per wave
        a_sweep = alg_sweep / (t_sweep * 1e-3) / 1e9
        r_sweep = {"bound": "hbm", "achieved": round(a_sweep, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(a_sweep / HBM_PEAK_GBS, 6), "traffic": None,
                   "kernel": "k_fm2d_wave<1,16>", "launch_solves": int(solves), "launch_ms": round(t_sweep, 3),
                   "alg_bytes": int(alg_sweep),
                   "note": "the reference's heap loop for every 2-D problem (the mixed-order times depend on "
                           "its pop order), four problems per wave with their heaps in LDS; algorithmic bytes "
                           "per solve 4 C + 4 Nr; bound by the pops' dependent latency and VALU issue, not HBM"}
    elif t_sweep > 0:
        a_sweep = alg_sweep / (t_sweep * 1e-3) / 1e9
        r_sweep = {"bound": "hbm", "achieved": round(a_sweep, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(a_sweep / HBM_PEAK_GBS, 6), "traffic": traffic_of("k_fm2d_strip<0>", n0),
                   "kernel": "k_fm2d_strip", "launch_solves": int(solves), "launch_ms": round(t_sweep, 3),
                   "alg_bytes": int(alg_sweep),
                   "note": "Gauss-Seidel sweeps of the coarse grid, one wave per (chain, period, source); "
                           "algorithmic bytes per solve 4 C + 4 Nr (the survey's 2-D definition without the "
                           "refined nodes, which the near-source kernel handles); the traffic counts every "
                           "sweep's re-read of the grid (the times and velocities are FP64 here); bound in "
                           "practice by FP64 VALU issue and latency, not HBM"}
    roof, other = r_disp, r_sweep
    if r_sweep and t_sweep > st_b[1]:
        roof, other = r_sweep, r_disp
    roof = dict(roof)
    roof["batch_parts"] = int(nparts)
    roof["stage_share"] = {k: round(v / max(sum(st_b), 1e-9), 4) for k, v in
                           zip(("nearest", "dispersion", "fm2d", "misfit"), st_b)}
    roof["fm2d_split_ms"] = {"near_source": round(t_near, 3), "sweep": round(t_sweep, 3)}
    if other:
        roof["next_kernel"] = other
    # whole evaluation against the survey's per-stage roofline:
    # t_roof = sum over stages of max(B_s / 8 TB/s, F_s / peak_s), per eval
    S_ = int(np.mean(N))
    R_ = (2 * 8 * 4 + 1) ** 2          # refined nodes, sgext 8, sgdic 4
    stages = {
        "nearest": (4.0 * G + 24.0 * S_, 8.0 * G * S_, FP64_PEAK_TFLOPS),
        "dispersion": (4.0 * G + 8.0 * C * npd, 170.0 * levB / B, FP64_PEAK_TFLOPS),
        "fm2d": (npd * nsrc * (4.0 * (C + R_) + 4.0 * nrev), npd * nsrc * 20.0 * (C + R_), FP32_PEAK_TFLOPS),
        "misfit": (20.0 * nsrc * nrev * npd, 8.0 * nsrc * nrev * npd, FP64_PEAK_TFLOPS),
    }
    t_roof = sum(max(b_ / (HBM_PEAK_GBS * 1e9), f_ / (pk * 1e12)) for b_, f_, pk in stages.values())
    t_eval = ms_step * 1e-3 / B
    roof["whole_eval"] = {
        "t_roof_ms": round(t_roof * 1e3, 4), "t_measured_ms": round(t_eval * 1e3, 4),
        "frac": round(t_roof / t_eval, 6),
        "hbm_only_frac": round(sum(v[0] for v in stages.values()) / t_eval / (HBM_PEAK_GBS * 1e9), 8),
        "stage_bound_ms": {k: round(max(b_ / (HBM_PEAK_GBS * 1e9), f_ / (pk * 1e12)) * 1e3, 5)
                           for k, (b_, f_, pk) in stages.items()},
        "note": "per evaluation: nearest B = 4G + 24S, F = 8GS (FP64); dispersion B = 4G + 8CP, F = 170 x "
                "dltar evaluations x (layers-1) of the sequential walk over every column (FP64); 2-D per "
                "solve B = 4(C+R) + 4Nr, F = 20(C+R) (FP32), R = 65^2; misfit B = 20 Ns Nr P, F = 8 Ns Nr P; "
                "S = the batch's mean cell count"}

    # ---- parity of this run vs the oracle, and the CPU baseline ------------------
    cpu = err = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        # every host core this process may run on (the GPU box's CPU share)
        threads = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        orc = po.surf_likelihood(gdict(w.grid), w.points, w.params, w.freqs, w.src, w.rev, w.ttime,
                                 w.raystat, order=args.order, nthreads=threads)
        t_cpu = time.perf_counter() - t0
        rc, out_g, pt_g, _, _, pv_g = like.likelihood(w.points, w.params, arrays=True)
        err = {"phase_velocity_max_abs_kms": float(np.max(np.abs(pv_g - orc[5]))),
               "traveltime_max_rel": float(np.max(np.abs(pt_g - orc[2]) / orc[2])),
               "loglike_rel": float(abs(out_g[0] - orc[1][0]) / abs(orc[1][0]))}
        cpu = {"value": round(1.0 / t_cpu, 5), "unit": "evals/s", "cores": threads, "kind": "port",
               "sample": "one full evaluation of the plain chain's model (nearest site, 40,401 "
                         "surfdisp96 columns, 640 modrays solves, misfit) by the C oracle, OpenMP "
                         "over columns and periods as the reference, %.1f s" % t_cpu}
        # the reference's own code where it builds here (oracle/_ref: surfdisp96 and
        # modrays compiled from the reference's Fortran with OpenMP over columns and
        # periods, as its OMP build runs them); nearest site, convert_to_layer and
        # the misfit are the oracle's (kdtree_to_grid and the likelihood loops need
        # the whole sampler to build)
        try:
            ref_ok = po.ref() is not None
        except OSError:
            ref_ok = False
        if ref_ok:
            def ref_eval(nth):
                t0 = time.perf_counter()
                rr = po.ref_surf_forward(gdict(w.grid), w.points, w.params, w.freqs, w.src, w.rev, w.raystat,
                                         order=args.order, nthreads=nth)
                po.body_misfit(rr[2], w.ttime, w.raystat, np.zeros((w.src.shape[0], 4)))
                return time.perf_counter() - t0, rr
            ref_eval(threads)                                   # warm-up
            runs = [ref_eval(threads) for _ in range(3)]
            t_ref = float(np.median([r[0] for r in runs]))
            rr = runs[-1][1]
            err["reference_stages_equal_oracle"] = bool(rr[0] == 0 and np.array_equal(rr[1], orc[5]) and
                                                        np.array_equal(rr[2], orc[2]))
            t_ref1 = ref_eval(1)[0] if not args.no_cpu_1thread else None
            cpu_model = "unknown"
            try:
                with open("/proc/cpuinfo") as f:
                    for line in f:
                        if line.startswith("model name"):
                            cpu_model = line.split(":", 1)[1].strip()
                            break
            except OSError:
                pass
            cpu = {"value": round(1.0 / t_ref, 5), "unit": "evals/s", "cores": threads, "kind": "reference",
                   "sample": "full evaluations of the plain chain's model: surfdisp96 on 40,401 columns "
                             "and modrays for 20 periods x 32 sources by the reference's own Fortran "
                             "(amdflang -O3 -fopenmp, OpenMP over columns and periods), nearest site, "
                             "convert_to_layer and misfit by the C oracle; median of 3 after a warm-up, "
                             "%.2f s on %d threads (the C oracle alone: %.2f s)" % (t_ref, threads, t_cpu),
                   "cpu_model": cpu_model, "host_cpus_visible": os.cpu_count(),
                   "one_thread_value": None if t_ref1 is None else round(1.0 / t_ref1, 5),
                   "one_thread_s": None if t_ref1 is None else round(t_ref1, 3),
                   "port_value": round(1.0 / t_cpu, 5)}
        log("cpu baseline", cpu, "err", err)

    return {
        "metric": METRIC, "value": round(batched_eps, 3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "surface-wave likelihood, 201x201x61 grid (100x100x30 km, 0.5 km), "
                               "Rayleigh phase velocity at 20 periods 1-10 s, 32 sources x 32 "
                               "receivers, order %d, refined source grid; batch of %d chains with "
                               "%d-%d Voronoi cells per step" % (args.order, B, cmin, cmax),
                   "batch": B, "grid": [w.grid.nx, w.grid.ny, w.grid.nz], "periods": npd,
                   "nsrc": w.nsrc, "nrev": int(w.rev.shape[0]), "cells": [cmin, cmax],
                   "plain_cells": int(w.points.shape[0]), "order": args.order,
                   "fm2d_method": ("heap loop, four problems per wave" if heap1 else
                                   "parallel sweeps + heap loop near the source") if args.fm2d_method == 0
                   else "sequential heap loop per (period, source)"},
        "plain_evals_per_s": round(plain_eps, 4),
        "p50_ms": round(float(np.median(lat)) * 1e3, 3),
        "p99_ms": round(float(np.quantile(lat, 0.99)) * 1e3, 3),
        "batched_vs_plain": round(batched_eps / plain_eps, 3),
        "plain_stage_ms": [round(v, 3) for v in st_plain],
        "served": served,
        "served_mix": mixres,
        "autotune": tuned,
        "perturbed_box": pbox,
        "group_velocity": group,
        "order1_batched": order1,
        "roofline": roof,
        "cpu_baseline": cpu,
        "max_error_vs_oracle": err,
    }


# --------------------------------------------------------------------- batched C3
def run_c3(args):
    """The batched configuration: B chains (default 256) per launch on a
    101x101x41 grid (50 x 50 x 20 km), cell counts uniform in [20,400], 20
    periods 0.5-10 s, 64 sources x 64 receivers, Rayleigh and Love, phase and
    group velocities.  One evaluation = one chain's four likelihoods, one
    context per (wave type, velocity type), each a batched launch of all
    chains.  The contexts run one after another (one holds its workspace at a
    time); a step is the four batches, and its time is the sum of theirs."""
    import mctomo_amd as mct
    from mctomo_amd.workloads import surface_workload, _surface_sites, top_is_slowest

    B = max(1, args.batch)
    cmin, cmax = 20, 400
    kinds = [(1, 0), (1, 1), (0, 0), (0, 1)]          # (raylov, phase_group)
    names = {(1, 0): "Rayleigh phase", (1, 1): "Rayleigh group", (0, 0): "Love phase", (0, 1): "Love group"}
    ws = {rl: surface_workload(101, 41, 50.0, 20.0, 200, 20, (0.5, 10.0), 64, 64, raylov=rl,
                               seed=20261204, order=args.order) for rl in (1, 0)}
    g = ws[1].grid
    L, Z = g.dx * (g.nx - 1), g.dz * (g.nz - 1)
    rng = np.random.default_rng(4243)
    P = np.zeros((B, cmax, 3))
    Q = np.zeros((B, cmax, 3))
    N = np.zeros(B, np.int32)
    for c in range(B):
        while True:       # check_model rejects a column whose top node is not its slowest
            k = int(rng.integers(cmin, cmax + 1))
            p, q = _surface_sites(rng, k, L, Z, (2.5, 4.5))
            if top_is_slowest(g, p, q):
                break
        P[c, :k], Q[c, :k], N[c] = p, q, k
    dP, dQ, dN = (mct.DeviceArray.from_numpy(a) for a in (P, Q, N))
    dS0 = mct.DeviceArray.from_numpy(np.zeros((B, 2, ws[1].freqs.shape[0])))
    dO = mct.DeviceArray((B, 3), np.float64)
    dS = mct.DeviceArray((B,), np.int32)
    per_kind = {}
    total_s = 0.0
    for rl, pg in kinds:
        w = ws[rl]
        like = mct.SurfLikelihood(w.grid, w.freqs, w.src, w.rev, w.ttime, w.raystat, raylov=rl,
                                  order=args.order, phase_group=pg, max_cells=cmax, max_chains=B)

        def run():
            like.likelihood_batch_dev(B, dP.ptr, dQ.ptr, dN.ptr, dS0.ptr, dO.ptr, dS.ptr)
            like.sync()

        for _ in range(max(1, args.warmup)):
            run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
        dt = (time.perf_counter() - t0) / args.steps
        st = dS.numpy()
        if (st < 0).any():
            raise RuntimeError("%s: chain status %s" % (names[(rl, pg)], st))
        stages = [round(float(v), 2) for v in like.timings()]
        near_ms, sweep_ms = like.timings2d()
        per_kind[names[(rl, pg)]] = {"ms_per_batch": round(dt * 1e3, 2), "stage_ms": stages,
                                     "fm2d_split_ms": {"near_source": round(near_ms, 2),
                                                       "sweep": round(sweep_ms, 2)},
                                     "statuses": {str(k): int(v) for k, v in
                                                  zip(*np.unique(st, return_counts=True))}}
        log("C3 %s: %.1f ms per %d-chain batch, stages(ms) nearest %.2f dispersion %.2f 2-D %.2f "
            "misfit %.2f (near-source %.2f, sweep %.2f)" % (names[(rl, pg)], dt * 1e3, B, *stages,
                                                           near_ms, sweep_ms))
        total_s += dt
        like.close()
    ms_step = total_s * 1e3
    value = B / total_s
    cpu = None
    if not args.no_cpu:
        # one chain's four likelihoods by the C oracle on every host core the process may use
        from oracle import pyoracle as po
        threads = len(os.sched_getaffinity(0))
        k = int(N[0])
        t0 = time.perf_counter()
        for rl, pg in kinds:
            w = ws[rl]
            po.surf_likelihood(gdict(w.grid), P[0, :k], Q[0, :k], w.freqs, w.src, w.rev, w.ttime,
                               w.raystat, raylov=rl, order=args.order, phase_group=pg, nthreads=threads)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(1.0 / t_cpu, 4), "unit": "evals/s", "cores": threads, "kind": "port",
               "sample": "one chain (%d cells), its four likelihoods by the C oracle, %.1f s" % (k, t_cpu)}
    return {
        "metric": METRIC, "value": round(value, 3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched surface-wave likelihood: 101x101x41 grid (50x50x20 km), 20 periods "
                               "0.5-10 s, 64 sources x 64 receivers, Rayleigh and Love, phase and group "
                               "velocities (four likelihoods per eval), order %d; batch of %d chains with "
                               "20-400 Voronoi cells" % (args.order, B),
                   "batch": B, "grid": [101, 101, 41], "periods": 20, "nsrc": 64, "nrev": 64,
                   "order": args.order},
        "per_kind": per_kind,
        "roofline": None,
        "cpu_baseline": cpu,
    }


# --------------------------------------------------------------------- body
def run_body(args):
    import mctomo_amd as mct
    from mctomo_amd.workloads import c2_body

    w = c2_body(order=args.order)
    B = max(1, args.batch)
    ncells = w.points.shape[0]
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, w.nsrc, order=args.order,
                              max_cells=ncells, max_chains=B)
    rng = np.random.default_rng(4242)
    L = w.grid.dx * (w.grid.nx - 1)
    P = rng.uniform(0.0, L, (B, ncells, 3))
    vp = rng.uniform(4.0, 7.0, (B, ncells))
    Q = np.stack([vp, vp / 1.73, 0.32 * vp + 0.77], axis=2)
    P[0], Q[0] = w.points, w.params
    loc = np.repeat(w.locations[None], B, axis=0)
    dP, dQ, dL = (mct.DeviceArray.from_numpy(a) for a in (P, Q, loc))
    dN = mct.DeviceArray.from_numpy(np.full(B, ncells, np.int32))
    dO = mct.DeviceArray((B, 3), np.float64)
    dS = mct.DeviceArray((B,), np.int32)

    def run(nchain):
        like.likelihood_batch_dev(nchain, dP.ptr, dQ.ptr, dN.ptr, dL.ptr, 0, dO.ptr, dS.ptr)
        like.sync()

    like.set_tuning(0.5)
    for _ in range(args.warmup):
        run(1)
    lat = []
    for _ in range(args.plain_evals):
        t0 = time.perf_counter()
        run(1)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat)
    plain_eps = 1.0 / lat.mean()
    like.set_tuning(0.25)
    for _ in range(args.warmup):
        run(B)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(B)
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    batched_eps = B * args.steps / dt
    like.set_kernel_timing(True)
    run(B)
    tb = like.timings()
    like.set_kernel_timing(False)
    G = w.grid.nx * w.grid.ny * w.grid.nz
    # the context's eikonal precision: FP64 velocities and local solves unless the
    # environment turns them off at create (times are FP64 either way)
    f64 = os.environ.get("MCT_EIK_F64", "1") != "0"
    alg = B * w.nsrc * (4.0 * G + 4.0 * w.rev.shape[0])
    heap = args.order == 1 and f64 and os.environ.get("MCT_EIK_ORDER1", "heap") != "parallel"
    if heap:
        # order 1 replays the reference's heap loop: one launch of k_sfmm<1> per batch
        # (one wave per source), timed by the context's stage events
        launches = 1
        avg_ms = tb["eikonal_ms"]
        kname = "k_sfmm<1>"
    else:
        launches = max(1, tb["round_launches"])
        avg_ms = tb["round_kernels_ms"] / launches
        kname = "k_eik_round<%d,double,%s>" % (args.order, "double" if f64 else "float")
    achieved = alg / launches / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None, "kernel": kname,
            "avg_launch_ms": round(avg_ms, 5), "launches": launches}
    if heap:
        roof["note"] = ("the heap replay is sequential per source (about 1.1 pops per cell, each a "
                        "chain of dependent memory round trips): memory latency, not HBM bandwidth, "
                        "bounds it; eikonal_ms includes the sign-bit pass")
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        threads = len(os.sched_getaffinity(0))
        grid = gdict(w.grid)
        t0 = time.perf_counter()
        v = po.kdtree_to_grid(grid, w.points, w.params)[1]
        pt = po.body_traveltimes(v, grid, w.locations, w.rev, args.order, threads)
        po.body_misfit(pt[None], w.ttime, w.raystat, w.locations)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": round(1.0 / t_cpu, 4), "unit": "evals/s", "cores": threads, "kind": "port",
               "sample": "one full evaluation by the C oracle, %.1f s" % t_cpu}
    return {
        "metric": METRIC, "value": round(batched_eps, 3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64" if f64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "3-D body-wave eikonal drop-in: 100 Voronoi sites, 101^3 at 0.2 km, "
                               "50 sources x 400 receivers, P, order %d%s; batch of %d" %
                               (args.order, " (the reference's heap loop replayed)" if heap else "", B),
                   "batch": B},
        "plain_evals_per_s": round(plain_eps, 3),
        "p50_ms": round(float(np.median(lat)) * 1e3, 3),
        "p99_ms": round(float(np.quantile(lat, 0.99)) * 1e3, 3),
        "batched_vs_plain": round(batched_eps / plain_eps, 3),
        "roofline": roof, "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=None,
                    help="chains per launch (default 64; 256 for --workload c3)")
    ap.add_argument("--order", type=int, default=0)
    ap.add_argument("--workload", default="surf201", choices=["surf201", "body101", "c3"])
    ap.add_argument("--fm2d-method", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-evals", type=int, default=10)
    ap.add_argument("--served", type=int, default=512, help="requests for the served measurement (0: skip)")
    ap.add_argument("--box-steps", type=int, default=3,
                    help="timed perturbed-box batches (secondary unit; 0: skip)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--served-mix", type=int, default=1024,
                    help="chains (one request each) of the three-grid served mix (0: skip)")
    ap.add_argument("--served-fill-ms", type=float, default=25.0,
                    help="longest wait of a partial batch while another lane runs (< 0: none)")
    ap.add_argument("--served-hbm-gb", type=int, default=200, help="device bytes the served lanes may hold")
    ap.add_argument("--group-steps", type=int, default=2,
                    help="timed group-velocity batches (secondary figure; 0: skip)")
    ap.add_argument("--order1-steps", type=int, default=3,
                    help="timed batches with the mixed-order update beside the order-0 value (0: skip)")
    ap.add_argument("--no-cpu-1thread", action="store_true", help="skip the one-thread CPU reference eval")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.batch is None:
        args.batch = 256 if args.workload == "c3" else 64
    if args.gpus != 1:
        log("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU "
            "path (chains would shard across GPUs with no device traffic); use --gpus 1.")
        return 2
    if args.calib_copy_mb > 0:
        # a device-to-device copy of known size: the reference reading for the
        # HBM traffic counters (FETCH_SIZE / WRITE_SIZE) of a profiled run
        import mctomo_amd as mct
        n = args.calib_copy_mb * (1 << 20) // 8
        a, b = mct.DeviceArray((n,), np.float64), mct.DeviceArray((n,), np.float64)
        mct.lib().mct_copy_d2d(ctypes_vp(b.ptr), ctypes_vp(a.ptr), ctypes_i64(8 * n))
        a.free(); b.free()
    run = {"surf201": run_surface, "body101": run_body, "c3": run_c3}[args.workload]
    res = run(args)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
