#!/usr/bin/env python3
"""bench.py -- likelihood evaluations per second on one MI355X.

Workload: the 3-D body-wave eikonal drop-in -- 100 Voronoi sites with Vp
uniform in [4,7] km/s rasterised on a 101x101x101 grid at 0.2 km, 50 sources
at 2-18 km depth, 400 surface receivers, first-arrival P times, first-order
fast marching.  One evaluation = one likelihood() call with the whole grid as
the perturbed box: nearest-site rasterisation, 50 eikonal solves, 20,000
picks, Gaussian misfit.

A step is one batch of B chains' evaluations (inputs already in HBM); the
plain single-request drop-in (B=1) is timed beside it for latency.  Prints one
JSON line.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_chains(base, nchain, ncells, seed):
    """Per-chain Voronoi models (own sites and Vp) on the base workload's data."""
    rng = np.random.default_rng(seed)
    L = base.grid.dx * (base.grid.nx - 1)
    P = rng.uniform(0.0, L, (nchain, ncells, 3))
    vp = rng.uniform(4.0, 7.0, (nchain, ncells))
    Q = np.stack([vp, vp / 1.73, 0.32 * vp + 0.77], axis=2)
    P[0], Q[0] = base.points, base.params
    loc = np.repeat(base.locations[None], nchain, axis=0)
    return P, Q, loc


def cpu_baseline(w, order, threads):
    """The oracle (C restatement of the reference path, OpenMP over sources as
    src/likelihood_body.F90:204-224) on the same model: one full evaluation on
    `threads` host threads."""
    from oracle import pyoracle as po
    g = w.grid
    grid = dict(nx=g.nx, ny=g.ny, nz=g.nz, xmin=g.xmin, ymin=g.ymin, zmin=g.zmin,
                dx=g.dx, dy=g.dy, dz=g.dz, scaling=1.0)
    t0 = time.perf_counter()
    vp = po.kdtree_to_grid(grid, w.points, w.params)[1]
    pt = po.body_traveltimes(vp, grid, w.locations, w.rev, order, threads)
    out, _, _, _ = po.body_misfit(pt[None], w.ttime, w.raystat, w.locations)
    dt = time.perf_counter() - t0
    return dt, pt, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--order", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-evals", type=int, default=20)
    ap.add_argument("--band-plain", type=float, default=0.5)
    ap.add_argument("--band-batched", type=float, default=0.25)
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU "
            "path (chains would shard across GPUs with no device traffic); use --gpus 1.")
        return 2

    import mctomo_amd as mct
    from mctomo_amd.workloads import c2_body

    w = c2_body(order=args.order)
    B = max(1, args.batch)
    ncells = w.points.shape[0]
    like = mct.BodyLikelihood(w.grid, w.rev, w.ttime, w.raystat, w.nsrc, order=args.order,
                              max_cells=ncells, max_chains=B)
    P, Q, loc = make_chains(w, B, ncells, 4242)
    dP, dQ, dL = (mct.DeviceArray.from_numpy(a) for a in (P, Q, loc))
    dN = mct.DeviceArray.from_numpy(np.full(B, ncells, np.int32))
    dO = mct.DeviceArray((B, 3), np.float64)
    dS = mct.DeviceArray((B,), np.int32)

    def run(nchain):
        like.likelihood_batch_dev(nchain, dP.ptr, dQ.ptr, dN.ptr, dL.ptr, 0, dO.ptr, dS.ptr)
        like.sync()

    # ---- plain single-request drop-in: latency per evaluation ----------------
    like.set_tuning(args.band_plain)
    for _ in range(args.warmup):
        run(1)
    lat = []
    for _ in range(args.plain_evals):
        t0 = time.perf_counter()
        run(1)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat)
    plain_eps = 1.0 / lat.mean()
    log("plain: mean %.3f ms p50 %.3f p99 %.3f" % (lat.mean() * 1e3, np.median(lat) * 1e3,
                                                   np.quantile(lat, 0.99) * 1e3), like.timings())

    # ---- batched: the timed region ------------------------------------------
    like.set_tuning(args.band_batched)
    for _ in range(args.warmup):
        run(B)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(B)
    dt = time.perf_counter() - t0
    status = dS.numpy()
    if (status != 0).any():
        log("non-zero chain status", status)
        return 1
    ms_step = dt / args.steps * 1e3
    batched_eps = B * args.steps / dt
    log("batched: B=%d %.3f ms/step %.1f evals/s" % (B, ms_step, batched_eps), like.timings())

    # ---- roofline of the dominant kernel (eikonal round), live HIP events -----
    like.set_kernel_timing(True)
    run(B)
    tb = like.timings()
    like.set_kernel_timing(False)
    G = w.grid.nx * w.grid.ny * w.grid.nz
    nsolve = w.nsrc * (1 if w.datatype == 0 else 2)
    alg_bytes_eval = nsolve * (4.0 * G + 4.0 * w.rev.shape[0])   # per-solve 4G + 4Nr bytes
    launches = max(1, tb["round_launches"])
    avg_launch_ms = tb["round_kernels_ms"] / launches
    alg_per_launch = B * alg_bytes_eval / launches
    achieved = alg_per_launch / (avg_launch_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": None,
            "kernel": "k_eik_round<0,double>", "avg_launch_ms": round(avg_launch_ms, 5),
            "launches": launches}

    # ---- parity of this run vs the oracle, and the CPU baseline --------------
    cpu = None
    err = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
        t_cpu, pt_o, out_o = cpu_baseline(w, args.order, threads)
        rc, out_g, pt_g, _, _ = like.likelihood(w.points, w.params, w.locations, arrays=True)
        err = {"traveltime_max_rel": float(np.max(np.abs(pt_g[0] - pt_o) / pt_o)),
               "loglike_rel": float(abs(out_g[0] - out_o[0]) / abs(out_o[0]))}
        cpu = {"value": round(1.0 / t_cpu, 4), "unit": "evals/s", "cores": threads,
               "kind": "port",
               "sample": "one full evaluation of this workload (nearest site + 50 SFMM solves "
                         "+ misfit) by the C oracle, OpenMP over sources, %.1f s" % t_cpu}
        log("cpu baseline", cpu, "err", err)

    res = {
        "metric": METRIC, "value": round(batched_eps, 3), "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "3-D body-wave eikonal drop-in: 100 Voronoi sites, Vp U[4,7] km/s, "
                               "101x101x101 grid at 0.2 km, 50 sources x 400 receivers, P, "
                               "order %d; batch of %d chains per step" % (args.order, B),
                   "batch": B, "grid": [w.grid.nx, w.grid.ny, w.grid.nz], "nsrc": w.nsrc,
                   "nrev": int(w.rev.shape[0]), "ncells": ncells, "order": args.order,
                   "precision": "eikonal local solve f32, arrival times f64; nearest-site, picks, "
                                "misfit f64", "band_tiles": [args.band_plain, args.band_batched]},
        "plain_evals_per_s": round(plain_eps, 3),
        "p50_ms": round(float(np.median(lat)) * 1e3, 3),
        "p99_ms": round(float(np.quantile(lat, 0.99)) * 1e3, 3),
        "batched_vs_plain": round(batched_eps / plain_eps, 3),
        "roofline": roof,
        "cpu_baseline": cpu,
        "max_error_vs_oracle": err,
    }
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
