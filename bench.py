"""Benchmark: VoroMesh optimizer iterations on one MI355X.

Default workload: 512 independent fits of uneven size (sites log-uniform in
[256, 65,536], Poisson-disk initialised; 8 targets per site on a random simple
polygon with 5e-4 jitter; k = 16; lr 1e-3), all advanced by one set of launches
per iteration (one iteration captured as a HIP graph).  A "step" is one
optimizer iteration of every instance.  Also measured: the same fits run one at
a time through the plain per-instance path, the loss kernel's duration (HIP
events on the batch's stream) for the roofline, and the CPU oracle on a bounded
sample of the same workload.

Prints one JSON line.  `value` = site-iterations per second of the batched form.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "bb-temp-testing_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "site-iters/s batched vs plain at <=1e-5 grad err vs FP64; p50/p99 fit latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def make_batched(n_inst, seed, lo=256, hi=65536, per_site=8):
    """Sites (Poisson, batched GPU sampler) and polygon targets for every instance."""
    import vmhip
    from vmhip import workloads as wl
    sizes = wl.log_uniform_sizes(n_inst, lo, hi, seed)
    radii = np.array([wl.poisson_radius_for(int(s)) for s in sizes], np.float32)
    seeds = (2 * (1 << 20) + np.arange(n_inst)).astype(np.uint32)
    grids, _, ms = vmhip.poisson_batch(np.ones(n_inst), np.ones(n_inst), radii, seeds)
    insts = []
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        tg = wl.polygon(per_site * len(sites), sigma=5e-4, seed=int(seeds[i]))
        insts.append((sites, tg))
    return insts, ms


def loss_bytes(S, T, C):
    """Algorithmic HBM bytes of one loss-kernel launch: targets read (8 B), sorted sites
    read (pos 8 B + index 4 B), cell starts read (4 B), gradient accumulated (8 B)."""
    return 8 * T + 20 * S + 4 * C


def iter_bytes(S, T, C):
    """Algorithmic bytes of one whole iteration: 8T + 88S + 8C."""
    return 8 * T + 88 * S + 8 * C


def run_batched(insts, K, steps, warmup, lr=1e-3):
    import vmhip
    b = vmhip.Batch()
    for sites, tg in insts:
        b.add(1.0, 1.0, sites, tg, lr, loss_k=K)
    t0 = time.time()
    b.commit()
    log(f"  batch committed in {time.time() - t0:.1f} s")
    if warmup:
        b.step(warmup)
    ms = b.step(steps)
    phase = b.profile(5)
    cells = 0
    for i in range(len(insts)):
        ch = b.choice(i)
        g = math.ceil(1.0 / ch["grid_cell"])
        cells += g * g + 1
    return b, ms, phase, cells


def run_plain(insts, K, iters, lr=1e-3):
    """The same fits one at a time: one plain per-instance engine per fit."""
    import vmhip
    total_ms, site_iters = 0.0, 0
    for sites, tg in insts:
        b = vmhip.Batch()
        b.add(1.0, 1.0, sites, tg, lr, loss_k=K)
        b.commit()
        b.step(1)
        total_ms += b.step(iters)
        site_iters += len(sites) * iters
        b.close()
    return site_iters / (total_ms / 1e3), total_ms


def run_serving(seconds, rate, seed=4, lo=512, hi=131072, budget=200e9, max_iters=300, tol=1e-6, K=16):
    """Poisson arrivals at `rate`/s for `seconds`; each request a fit of log-uniform size with 8
    targets per site on a random polygon, run to |dL/L| < tol or max_iters; 200 GB budget."""
    import vmhip
    from vmhip import workloads as wl
    arr = wl.poisson_arrivals(rate, seconds, seed)
    n = len(arr)
    sizes = wl.log_uniform_sizes(n, lo, hi, seed + 1)
    radii = np.array([wl.poisson_radius_for(int(s)) for s in sizes], np.float32)
    seeds = (4 * (1 << 20) + np.arange(n)).astype(np.uint32)
    grids, _, _ = vmhip.poisson_batch(np.ones(n), np.ones(n), radii, seeds)
    reqs = []
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        reqs.append(dict(width=1.0, height=1.0, sites=sites, targets=wl.polygon(8 * len(sites), sigma=5e-4,
                                                                              seed=int(seeds[i])),
                         lr=1e-3, loss_k=K))
    res, st = vmhip.serve(reqs, arr, hbm_budget_bytes=budget, max_iterations=max_iters, tol=tol, poll_every=10)
    lat = np.array([r["latency_s"] for r in res])
    its = np.array([r["iterations"] for r in res])
    return dict(requests=n, arrival_rate=rate, window_s=seconds, wall_s=st["wall_s"],
                p50_latency_s=float(np.percentile(lat, 50)), p99_latency_s=float(np.percentile(lat, 99)),
                requests_per_s=n / st["wall_s"], site_iters_per_s=st["site_iters"] / st["wall_s"],
                mean_iterations=float(its.mean()), max_resident=st["max_resident_seen"],
                peak_resident_GB=st["peak_bytes"] / 1e9, sites_total=int(sum(len(r["sites"]) for r in reqs)))


def cpu_baseline(insts, K, iters=10, budget_sites=500_000, lr=1e-3):
    import oracle as O
    threads = O.num_threads()
    sample, n = [], 0
    for s, t in insts:
        if n >= budget_sites:
            break
        sample.append((s, t))
        n += len(s)
    sc = O.make_sched(lr, 1.0, 1.0)
    t0 = time.perf_counter()
    si = 0
    for s, t in sample:
        cell = 2.0 / math.sqrt(len(s))
        O.fit(s, t, K, cell, sc, iters)
        si += len(s) * iters
    dt = time.perf_counter() - t0
    return dict(value=si / dt, unit="site-iters/s", cores=threads, kind="port",
                sample=f"first {len(sample)} of {len(insts)} instances ({n} sites, {sum(len(t) for _, t in sample)} "
                       f"targets), {iters} iterations each, oracle restatement (OpenMP {threads} threads, "
                       f"{cpu_model()})")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--plain-iters", type=int, default=20)
    ap.add_argument("--plain-instances", type=int, default=0, help="0: all instances")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--serve-seconds", type=float, default=0.25, help="arrival window of the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is a design "
              "note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import vmhip
    assert vmhip.device_count() > 0, "no HIP device"
    log(f"generating {args.instances} instances ...")
    t0 = time.time()
    insts, poisson_ms = make_batched(args.instances, args.seed)
    S = sum(len(s) for s, _ in insts)
    T = sum(len(t) for _, t in insts)
    log(f"  {S} sites, {T} targets in {time.time() - t0:.1f} s (Poisson {poisson_ms:.1f} ms GPU)")

    log(f"batched: warmup {args.warmup}, timed {args.steps} iterations ...")
    b, ms, phase, C = run_batched(insts, args.k, args.steps, args.warmup)
    value = S * args.steps / (ms / 1e3)
    ms_per_step = ms / args.steps
    log(f"  {ms_per_step:.3f} ms/iteration, {value:.3e} site-iters/s; phases ms {np.round(phase, 4).tolist()}")

    npl = args.plain_instances or len(insts)
    log(f"plain: {npl} fits one at a time, {args.plain_iters} iterations each ...")
    plain_rate, plain_ms = run_plain(insts[:npl], args.k, args.plain_iters)
    log(f"  plain {plain_rate:.3e} site-iters/s")

    lb = loss_bytes(S, T, C)
    loss_ms = float(phase[2])
    achieved = lb / (loss_ms / 1e3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "site-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Poisson-disk sites, random-polygon targets)",
        "config": {"workload": "512-instance mixed batch: sites log-uniform [256, 65536], 8 targets/site on a "
                               "random polygon (sigma 5e-4), k=16, lr 1e-3, one persistent set of launches per "
                               "iteration (HIP graph)",
                   "instances": len(insts), "sites": S, "targets": T, "hash_cells": C, "k": args.k},
        "plain_site_iters_per_s": plain_rate,
        "batched_over_plain": value / plain_rate,
        "phase_ms": {"scan": float(phase[0]), "scatter": float(phase[1]), "loss": float(phase[2]),
                     "adam_bin": float(phase[3])},
        "iteration_roofline": {"bytes": iter_bytes(S, T, C),
                               "achieved_GBs": iter_bytes(S, T, C) / (ms_per_step / 1e3) / 1e9},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "lossKernel<16,float>",
                     "algorithmic_bytes": lb, "kernel_ms": loss_ms},
    }
    if args.serve_seconds > 0:
        log(f"serving: Poisson arrivals {args.serve_rate}/s for {args.serve_seconds} s ...")
        sv = run_serving(args.serve_seconds, args.serve_rate)
        sv["vs_plain_one_at_a_time"] = sv["site_iters_per_s"] / plain_rate
        out["serving"] = sv
        log(f"  {sv['requests']} requests, p50 {sv['p50_latency_s']:.3f} s, p99 {sv['p99_latency_s']:.3f} s, "
            f"{sv['site_iters_per_s']:.3e} site-iters/s")
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(insts, args.k)
        log(f"  cpu {out['cpu_baseline']['value']:.3e} site-iters/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
