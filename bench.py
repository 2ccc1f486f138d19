"""Benchmark: VoroMesh optimizer iterations on one MI355X.

Default workload: 512 independent fits of uneven size (sites log-uniform in
[256, 65,536], Poisson-disk initialised; 8 targets per site on a random simple
polygon with 5e-4 jitter; k = 16; lr 1e-3), all advanced by one set of launches
per iteration (one iteration captured as a HIP graph).  A "step" is one
optimizer iteration of every instance.  Also measured: the same fits run one at
a time through the plain per-instance path, the loss kernel's duration (HIP
events on the batch's stream) for the roofline, and the CPU oracle on a bounded
sample of the same workload (the oracle is used only for those CPU baselines).

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


def measured_traffic(key):
    """HBM bytes per loss-kernel launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    (profiles/loss_kernel_traffic.json, calibrated on a 256 MiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "loss_kernel_traffic.json")
    try:
        with open(path) as f:
            return float(json.load(f)[key]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


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
    b.paths = b.loss_stats()
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


def run_large(iters=500, K=16, n_sites=262144, n_targets=4194304):
    """One large instance: Poisson sites (~262k) and 4.2M targets on 64 ellipses (sigma 2e-4)."""
    import vmhip
    from vmhip import workloads as wl
    r = wl.poisson_radius_for(n_sites)
    grids, _, _ = vmhip.poisson_batch([1.0], [1.0], [r], [1])
    sites = wl.live_sites(grids[0])
    tg = wl.ellipses(n_targets, count=64, sigma=2e-4, seed=1)
    b = vmhip.Batch()
    b.add(1.0, 1.0, sites, tg, 1e-3, loss_k=K)
    b.commit()
    stats0 = b.loss_stats()
    b.step(2)
    ms = b.step(iters)
    phase = b.profile(10)
    stats1 = b.loss_stats()
    ch = b.choice(0)
    g = math.ceil(1.0 / ch["grid_cell"])
    S, T, C = len(sites), len(tg), g * g + 1
    lb = loss_bytes(S, T, C)
    ib = iter_bytes(S, T, C)
    return dict(sites=S, targets=T, iterations=iters, ms_per_iter=ms / iters,
                site_iters_per_s=S * iters / (ms / 1e3),
                phase_ms=dict(zip(["scan", "scatter", "loss", "exact", "adam_bin"], [float(x) for x in phase])),
                loss_kernel_GBs=lb / (phase[2] / 1e3) / 1e9, loss_kernel_frac=lb / (phase[2] / 1e3) / 1e9 / HBM_PEAK_GBS,
                iteration_GBs=ib / (ms / iters / 1e3) / 1e9,
                iteration_frac=ib / (ms / iters / 1e3) / 1e9 / HBM_PEAK_GBS, grid_cell=ch["grid_cell"],
                loss_paths_start=stats0, loss_paths_end=stats1, loss_kernel_traffic_bytes=measured_traffic("large262k"),
                loss_kernel_algorithmic_bytes=lb,
                final_mean_loss=b.status(0)["last_mean_loss"], fp32_vs_fp64_gpu=large_precision(grids[0], r, tg, K))


def large_precision(grid, r, tg, K):
    """One loss pass of the large instance in FP32 and in the all-FP64 mode: norm-wise, max and mean gradient
    error, and how many targets select a different nearest site or bisector partner (near-ties below FP32
    resolution, whose terms dominate the norm-wise figure)."""
    import vmhip
    gw, gh, cs = vmhip.poisson_grid_dims(1.0, 1.0, r)
    lp = {}
    for prec in (0, 1):
        o = vmhip.Optimizer(1.0, 1.0, r, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / r, precision=prec)
        o.set_targets(tg)
        o.set_generators(grid, gw, gh, cs)
        lp[prec] = o.loss_pass(len(tg), want_terms=True)
        o.close()
    g32, g64 = lp[0]["grad"], lp[1]["grad"]
    ge = np.abs(g32 - g64)
    flips = int(np.count_nonzero((lp[0]["nearest"] != lp[1]["nearest"]) | (lp[0]["partner"] != lp[1]["partner"])))
    return dict(grad_rel_norm_err=float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
                grad_max_abs_err=float(ge.max()), grad_mean_abs_err=float(ge.mean()), selection_flips=flips)


def run_precision(iters=5, K=16):
    """Jittered 256x256 lattice (+-1e-7) and 524,288 circle targets: the FP32 run (FP64 clip and loss sum)
    against the all-FP64 GPU run: gradient, positions after `iters` steps, clipped-cell topology.  (Both runs
    against the FP64 CPU oracle are tests: tests/test_gpu_parity.py.)"""
    import vmhip
    from vmhip import workloads as wl
    sites = wl.lattice(256, 1e-7, seed=3)
    tg = wl.circle(524288, radius=0.3, sigma=1e-4, seed=3)
    runs = {}
    for prec in (0, 1):
        opt = vmhip.Optimizer(1.0, 1.0, 1 / 256, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 * 256,
                              iterations=iters, precision=prec)
        opt.set_targets(tg)
        opt.set_generators(sites, 256, 256, 1 / 256)
        grad = opt.loss_pass(len(tg), want_terms=False)["grad"]
        mesh0 = opt.extract()
        opt.step(iters)
        runs[prec] = (grad, mesh0, opt.generators().astype(np.float64))
    g64, m64, p64 = runs[1]
    grad, mesh0, pos = runs[0]
    ge, pe = np.abs(grad - g64), np.abs(pos - p64)
    mism = 0
    for i in range(len(sites)):
        a = mesh0["cell_labels"][mesh0["cell_start"][i]:mesh0["cell_start"][i + 1]]
        b = m64["cell_labels"][m64["cell_start"][i]:m64["cell_start"][i + 1]]
        mism += int(not np.array_equal(a, b))
    fp32 = dict(grad_rel_norm_err=float(np.linalg.norm(grad - g64) / np.linalg.norm(g64)),
                grad_max_abs_err=float(ge.max()), grad_mean_abs_err=float(ge.mean()),
                pos_max_abs_err=float(pe.max()), pos_mean_abs_err=float(pe.mean()),
                topology_mismatches=mism, unsafe_cells=mesh0["unsafe_cells"], overflow_cells=mesh0["overflow_cells"])
    return dict(fp32_vs_fp64_gpu=fp32, fp64_unsafe_cells=m64["unsafe_cells"], fp64_overflow_cells=m64["overflow_cells"],
                sites=len(sites), targets=len(tg), iterations=iters)


def run_local_edit(timeout=120):
    """The demo's scripted edit sequence, headless (tests/cpp/shot_demo, apps/voromesh_demo.cpp:526-570):
    per-stamp latency of the local pipeline (canvas, upload, Poisson resample, 100 local iterations,
    extract) on a 1024^2 canvas with Poisson r = 24 generators.  Run as a child process."""
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "shot_demo")
    if not os.path.exists(exe):
        return {"skipped": "tests/cpp/shot_demo not built"}
    r = subprocess.run([exe, "24", "1024", "100", "--json"], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"shot_demo failed ({r.returncode}): {r.stderr[-500:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def run_small(iters=200, K=16):
    """Plain single fit: Poisson r=0.025 seed 0, 8,192 star targets (sigma 1e-3), 200 iterations, clip;
    its CPU baseline is the oracle restatement over the same fit (cpu_baseline_small)."""
    import vmhip
    from vmhip import workloads as wl
    r = 0.025
    smp = vmhip.Sampler(1.0, 1.0, r, 30, 0)
    grid, gw, gh, cs = smp.cells(), smp.grid_w, smp.grid_h, smp.cell_size
    tg = wl.star(8192, sigma=1e-3, seed=0)
    sites = wl.live_sites(grid)
    opt = vmhip.Optimizer(1.0, 1.0, r, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / r, iterations=iters,
                          grid_cell=2.0 / math.sqrt(len(sites)))
    opt.set_targets(tg)
    opt.set_generators(grid, gw, gh, cs)
    t0 = time.perf_counter()
    gpu_ms = opt.optimize()
    mesh = opt.extract()
    wall = time.perf_counter() - t0
    out = dict(sites=len(sites), targets=len(tg), iterations=iters, gpu_ms=gpu_ms,
               gpu_site_iters_per_s=len(sites) * iters / (gpu_ms / 1e3), extract_gpu_ms=mesh["gpu_ms"],
               extract_host_ms=mesh["host_ms"], wall_s=wall, watertight=mesh["watertight"],
               final_mean_loss=opt.evaluate_loss())
    out.update(cpu_baseline_small(sites, tg, r, gw, gh, cs, K, iters))
    return out


def cpu_baseline_small(sites, tg, r, gw, gh, cs, K, iters):
    """The oracle restatement timed over the same plain fit (200 iterations and the clip), on the host's cores."""
    import oracle as O
    lr = float(np.float32(np.float32(1e-3 / r) * np.float32(r)))
    sc = O.make_sched(lr, 1.0, 1.0)
    m = np.float32(2) * np.float32(r)
    box = (float(-m), float(-m), float(np.float32(gw) * np.float32(cs) + m), float(np.float32(gh) * np.float32(cs) + m))
    t0 = time.perf_counter()
    ref, _, _, _, _ = O.fit(sites, tg, K, 2.0 / math.sqrt(len(sites)), sc, iters, box=box)
    O.clip(ref, 1.0, 1.0, 2.0 / math.sqrt(len(sites)))
    cpu_s = time.perf_counter() - t0
    return dict(cpu_s=cpu_s, cpu_threads=O.num_threads(), cpu_site_iters_per_s=len(sites) * iters / cpu_s)


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
    ap.add_argument("--no-extra", action="store_true", help="skip the large/small/precision sections")
    ap.add_argument("--only", choices=["large", "small", "precision", "local_edit"], default=None,
                    help="run one section alone and print its JSON (profiling)")
    ap.add_argument("--large-iters", type=int, default=500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is a design "
              "note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import vmhip
    assert vmhip.device_count() > 0, "no HIP device"
    # A 256 MiB device copy: the known-size kernel the HBM traffic counters are calibrated on.
    copy_ms = vmhip.copy_probe(256 << 20)
    log(f"copy probe: 256 MiB in {copy_ms:.3f} ms ({2 * 256 * 1.048576 / copy_ms:.0f} GB/s read+write)")
    if args.only:
        fn = {"large": lambda: run_large(iters=args.large_iters), "small": run_small, "precision": run_precision,
              "local_edit": run_local_edit}
        print(json.dumps(fn[args.only](), default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
        return
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
        "loss_paths": b.paths,
        "phase_ms": {"scan": float(phase[0]), "scatter": float(phase[1]), "loss": float(phase[2]),
                     "exact": float(phase[3]), "adam_bin": float(phase[4])},
        "iteration_roofline": {"bytes": iter_bytes(S, T, C),
                               "achieved_GBs": iter_bytes(S, T, C) / (ms_per_step / 1e3) / 1e9},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("batch512") if
                     args.instances == 512 and args.seed == 2 else None, "kernel": "lossKernel<16,float,false>",
                     "algorithmic_bytes": lb, "kernel_ms": loss_ms},
    }
    if not args.no_extra:
        log("config 1: large single instance ...")
        out["large_instance"] = run_large(iters=args.large_iters)
        li = out["large_instance"]
        log(f"  {li['ms_per_iter']:.3f} ms/iter, {li['site_iters_per_s']:.3e} site-iters/s, loss kernel "
            f"{li['loss_kernel_GBs']:.1f} GB/s")
        log("config 0: small plain fit (GPU and CPU oracle) ...")
        out["small"] = run_small()
        log(f"  {out['small']}")
        log("config 3: precision stress ...")
        out["precision"] = run_precision()
        log(f"  {out['precision']}")
        log("local edits: headless scripted demo sequence ...")
        out["local_edit"] = run_local_edit()
        le = out["local_edit"]
        if "total_ms" in le:
            log(f"  {le['stamps']} stamps, total {le['total_ms']['mean']:.3f} ms mean per edit "
                f"(optimize {le['optimize_ms']['mean']:.3f}, poisson {le['poisson_ms']['mean']:.3f}, "
                f"extract {le['extract_ms']['mean']:.3f}); watertight {le['after']['watertight']}")
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
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)


if __name__ == "__main__":
    main()
