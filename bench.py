"""Benchmark: VoroMesh optimizer iterations on one MI355X.

Default workload (`value`): 512 independent fits of uneven size (sites log-uniform in [256, 65,536], Poisson-disk
initialised; 8 targets per site on a random simple polygon with 5e-4 jitter; lr 1e-3; k, grid cell, chunk extent and
targets per lane from the tuner's table per size class), all advanced by one set of launches per iteration (one
iteration captured as a HIP graph).  A "step" is one optimizer iteration of every instance.

Also measured, each with its own section of the JSON line:
  * the same fits one at a time through the plain per-instance path, and the batch with k fixed at 16;
  * the gradient gate: FP32 gradient against the all-FP64 GPU run evaluated with the FP32 run's discrete choices
    (nearest site, bisector partner), selection flips counted separately -- for the batch, the large instance and
    the small fit;
  * the large single instance (262,144-site class, 4,194,304 targets on 64 ellipses, 500 iterations, final clip),
    with HBM GB/s for every kernel of an iteration;
  * the small plain fit (1,024-site class, star, 200 iterations) and the precision stress (256^2 lattice);
  * the boundary mesh of each with occupancy rasterized from its target shape;
  * local edits (the headless scripted demo), and the serving loop;
  * the CPU restatement (oracle) timed on the host's cores on bounded samples: the CPU baseline.
The oracle is used only by the CPU-baseline legs.

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
GRAD_GATE = 1e-5
PHASES = ["scan", "scatter", "loss", "exact", "adam_bin"]


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


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------
def make_batched(n_inst, seed):
    """Sites (Poisson, batched GPU sampler) and polygon targets for every instance."""
    import vmhip
    from vmhip import workloads as wl
    radii, seeds = wl.batch_plan(n_inst, seed)
    grids, _, ms = vmhip.poisson_batch(np.ones(n_inst), np.ones(n_inst), radii, seeds)
    insts = []
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        insts.append((sites, wl.batch_targets(sites, seeds[i])))
    return insts, seeds, ms


def occupancy(outlines, raster=1024):
    """Occupancy masks of the union of closed outlines (disjoint shapes) in the unit square, rasterized by the
    canvas (even-odd through voxel centres): (masks, tiles_x, tiles_y, voxel)."""
    import vmhip
    masks = None
    for poly in outlines:
        c = vmhip.Canvas(1.0, 1.0, 1.0 / raster)
        c.rasterize(np.asarray(poly, np.float32))
        m = c.inside_masks()
        masks = m if masks is None else masks | m
        tx, ty, vox = c.tiles_x, c.tiles_y, c.voxel_size
        c.close()
    return masks, tx, ty, vox


def mesh_summary(m):
    return dict(cells=int(len(m["cell_start"]) - 1), inside=int(np.count_nonzero(m["inside"])),
                boundary_vertices=int(len(m["vertices"])), boundary_edges=int(len(m["edges"])),
                watertight=bool(m["watertight"]), loops=int(m["loops"]), unsafe_cells=int(m["unsafe_cells"]),
                overflow_cells=int(m["overflow_cells"]), gpu_ms=float(m["gpu_ms"]), host_ms=float(m["host_ms"]))


# ---------------------------------------------------------------------------
# Algorithmic bytes
# ---------------------------------------------------------------------------
def loss_bytes(S, T, C):
    """One loss-kernel launch: targets read (8 B), sorted sites read (pos 8 B + index 4 B), cell starts read
    (4 B), gradient accumulated (8 B per site)."""
    return 8 * T + 20 * S + 4 * C


def iter_bytes(S, T, C):
    """One whole iteration: 8T + 88S + 8C."""
    return 8 * T + 88 * S + 8 * C


def phase_bytes(S, T, C):
    """Algorithmic bytes of each kernel of an iteration: scan (cell count read, start written, count cleared:
    12 B/cell), scatter (position 8, cell and rank 8, cell start 4 read; sorted position 8 and index 4 written:
    32 B/site), loss (above), exact selection (data-dependent, not counted), Adam + bin (position, m, v read and
    written 48, gradient read and cleared 32, cell and rank written 8: 88 B/site)."""
    return dict(scan=12 * C, scatter=32 * S, loss=loss_bytes(S, T, C), exact=None, adam_bin=88 * S)


def phase_report(phase_ms, S, T, C):
    out = {}
    for name, ms, b in zip(PHASES, phase_ms, [phase_bytes(S, T, C)[p] for p in PHASES]):
        d = dict(ms=float(ms))
        if b is not None and ms > 0:
            d.update(bytes=int(b), GBs=b / (ms / 1e3) / 1e9, frac=b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS)
        out[name] = d
    return out


# ---------------------------------------------------------------------------
# Gradient gate: FP32 against FP64 with the FP32 run's discrete choices
# ---------------------------------------------------------------------------
def term_grads64(sites, targets, t_idx, c0, cj):
    """FP64 gradient contributions of the VoroLoss terms of targets t_idx with nearest site c0 and bisector partner cj
    (voro_optimizer.cu:244-254): with u = p - c0, e = cj - c0, l = |e|, n = e / l, d = <u, e> / l - l / 2,
    dL/dcj = 2 d ((u - <u, n> n) / l - n / 2) and dL/dc0 = -2 d n - dL/dcj."""
    p = targets[t_idx].astype(np.float64)
    a = sites[c0].astype(np.float64)
    b = sites[cj].astype(np.float64)
    u, e = p - a, b - a
    ln = np.hypot(e[:, 0], e[:, 1])[:, None]
    n = e / ln
    un = np.sum(u * n, axis=1)[:, None]
    d = np.sum(u * e, axis=1)[:, None] / ln - 0.5 * ln
    dcj = (u - un * n) / ln - 0.5 * n
    gj = 2 * d * dcj
    g0 = -2 * d * n - gj
    return g0, gj


def gradient_gate(sites_list, targets_list, p32, p64):
    """p32, p64: loss passes (dicts with grad [sum S, 2], nearest/partner [sum T]) over the same fits at the same sites
    in FP32 and in the all-FP64 mode.  The FP64 gradient is re-evaluated with the FP32 run's choices where they
    differ (flips: targets whose nearest site or partner differs), then compared norm-wise and per site."""
    g32, g64 = p32["grad"].copy(), p64["grad"].copy()
    n32, j32, n64, j64 = p32["nearest"], p32["partner"], p64["nearest"], p64["partner"]
    flip = (n32 != n64) | (j32 != j64)
    s_off = np.cumsum([0] + [len(s) for s in sites_list])
    t_off = np.cumsum([0] + [len(t) for t in targets_list])
    for i in np.unique(np.searchsorted(t_off, np.nonzero(flip)[0], side="right") - 1):
        idx = np.nonzero(flip[t_off[i]:t_off[i + 1]])[0]
        sites, tg = sites_list[i], targets_list[i]
        for (near, part, sign) in ((n64, j64, -1.0), (n32, j32, 1.0)):
            nn, pp = near[t_off[i] + idx], part[t_off[i] + idx]
            ok = (nn >= 0) & (pp >= 0)
            if not ok.any():
                continue
            g0, gj = term_grads64(sites, tg, idx[ok], nn[ok], pp[ok])
            np.add.at(g64, s_off[i] + nn[ok], sign * g0)
            np.add.at(g64, s_off[i] + pp[ok], sign * gj)
    err = np.abs(g32 - g64)
    rel = float(np.linalg.norm(g32 - g64) / max(np.linalg.norm(g64), 1e-300))
    out = dict(grad_rel_norm_err=rel, grad_max_abs_err=float(err.max()), grad_mean_abs_err=float(err.mean()),
               grad_max_abs=float(np.abs(g64).max()), selection_flips=int(flip.sum()), targets=int(len(flip)),
               tolerance=GRAD_GATE, passed=bool(rel <= GRAD_GATE))
    if t_off[-1] <= 5_000_000:
        # Conditioning: a site's gradient is a sum of terms that cancel near convergence, and a sum in FP32 can
        # only be as exact as eps times the sum of its terms' magnitudes.  Per site and component, the error
        # over that magnitude (FP64 terms with the FP32 run's choices).
        mass = np.zeros_like(g64)
        for i in range(len(sites_list)):
            nn, pp = n32[t_off[i]:t_off[i + 1]], j32[t_off[i]:t_off[i + 1]]
            idx = np.nonzero((nn >= 0) & (pp >= 0))[0]
            if len(idx):
                g0, gj = term_grads64(sites_list[i], targets_list[i], idx, nn[idx], pp[idx])
                np.add.at(mass, s_off[i] + nn[idx], np.abs(g0))
                np.add.at(mass, s_off[i] + pp[idx], np.abs(gj))
        has = mass > 0
        ratio = err[has] / mass[has]
        out.update(grad_err_over_term_mass_max=float(ratio.max()) if ratio.size else 0.0,
                   grad_err_over_term_mass_mean=float(ratio.mean()) if ratio.size else 0.0,
                   grad_rel_norm_to_term_mass=float(np.linalg.norm(g32 - g64) / max(np.linalg.norm(mass), 1e-300)))
    return out


# ---------------------------------------------------------------------------
# Sections
# ---------------------------------------------------------------------------
def run_batched(insts, steps, warmup, loss_k=0, lr=1e-3):
    import vmhip
    b = vmhip.Batch()
    for sites, tg in insts:
        b.add(1.0, 1.0, sites, tg, lr, loss_k=loss_k)
    t0 = time.time()
    b.commit()
    log(f"  batch committed in {time.time() - t0:.1f} s")
    if warmup:
        b.step(warmup)
    ms = b.step(steps)
    return b, ms


def batch_cells(b, n):
    cells = 0
    for i in range(n):
        g = math.ceil(1.0 / b.choice(i)["grid_cell"])
        cells += g * g + 1
    return cells


def batch_gate(b, insts):
    """Gradient gate of the batch at its current sites: FP32 loss pass, then an all-FP64 batch of the same fits
    (same sites, same k) and its loss pass."""
    import vmhip
    sites = [b.sites(i) for i in range(len(insts))]
    tgs = [t for _, t in insts]
    p32 = b.loss_pass()
    b64 = vmhip.Batch(precision=1)
    for i, s in enumerate(sites):
        b64.add(1.0, 1.0, s, tgs[i], 1e-3, loss_k=b.choice(i)["loss_k"])
    b64.commit()
    p64 = b64.loss_pass()
    b64.close()
    return gradient_gate(sites, tgs, p32, p64)


def run_plain(insts, iters, lr=1e-3):
    """The same fits one at a time: one plain per-instance engine per fit (the tuner's choices, as in the batch)."""
    import vmhip
    total_ms, site_iters = 0.0, 0
    for sites, tg in insts:
        b = vmhip.Batch()
        b.add(1.0, 1.0, sites, tg, lr, loss_k=0)
        b.commit()
        b.step(1)
        total_ms += b.step(iters)
        site_iters += len(sites) * iters
        b.close()
    return site_iters / (total_ms / 1e3), total_ms


def batch_meshes(b, insts, seeds, count=4):
    """Boundary meshes of a few batch fits (smallest, largest and two between) with occupancy from their polygons."""
    from vmhip import workloads as wl
    order = np.argsort([len(s) for s, _ in insts])
    pick = [int(order[int(q * (len(order) - 1))]) for q in np.linspace(0, 1, count)]
    out = []
    for i in pick:
        masks, tx, ty, vox = occupancy([wl.polygon_outline(int(seeds[i]))])
        m = b.extract(i, masks, (tx, ty), vox)
        out.append(dict(instance=i, sites=len(insts[i][0]), **mesh_summary(m)))
    return out


def single_gate(sites_grid, gw, gh, cs, r, tg, K, grid_cell=0.0):
    """Gradient gate of one optimizer's fit at the given slots: FP32 and all-FP64 loss passes."""
    import vmhip
    from vmhip import workloads as wl
    lp = {}
    for prec in (0, 1):
        o = vmhip.Optimizer(1.0, 1.0, r, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / r, precision=prec,
                            grid_cell=grid_cell)
        o.set_targets(tg)
        o.set_generators(sites_grid, gw, gh, cs)
        lp[prec] = o.loss_pass(len(tg))
        o.close()
    live = np.nonzero(sites_grid[:, 0] >= 0)[0]
    site_of = -np.ones(len(sites_grid), np.int64)
    site_of[live] = np.arange(len(live))
    conv = {}
    for prec in (0, 1):
        d = lp[prec]
        conv[prec] = dict(grad=d["grad"][live], nearest=np.where(d["nearest"] >= 0, site_of[np.maximum(d["nearest"], 0)], -1),
                          partner=np.where(d["partner"] >= 0, site_of[np.maximum(d["partner"], 0)], -1))
    return gradient_gate([wl.live_sites(sites_grid)], [tg], conv[0], conv[1])


def run_large(iters=500, K=16, n_sites=262144, n_targets=4194304):
    """One large instance: Poisson sites (~262k) and 4.2M targets on 64 ellipses (sigma 2e-4), k = 16 as the workload
    states; 500 iterations; per-kernel GB/s; final clip with the ellipses' occupancy; gradient gate."""
    import vmhip
    from vmhip import workloads as wl
    r = wl.poisson_radius_for(n_sites)
    gw, gh, cs = vmhip.poisson_grid_dims(1.0, 1.0, r)
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
    masks, tx, ty, vox = occupancy(wl.ellipse_outlines(64, seed=1))
    t0 = time.perf_counter()
    mesh = b.extract(0, masks, (tx, ty), vox)
    extract_wall = (time.perf_counter() - t0) * 1e3
    fitted = b.sites(0)
    saturated = b.saturation()
    b.close()
    # gate at the fitted sites (a slot grid of the live sites)
    gate = single_gate(fitted, len(fitted), 1, 1.0 / len(fitted), r, tg, K)
    lb, ib = loss_bytes(S, T, C), iter_bytes(S, T, C)
    return dict(sites=S, targets=T, hash_cells=C, iterations=iters, ms_per_iter=ms / iters,
                site_iters_per_s=S * iters / (ms / 1e3), kernels=phase_report(phase, S, T, C),
                loss_kernel_GBs=lb / (phase[2] / 1e3) / 1e9, loss_kernel_frac=lb / (phase[2] / 1e3) / 1e9 / HBM_PEAK_GBS,
                iteration_bytes=ib, iteration_GBs=ib / (ms / iters / 1e3) / 1e9,
                iteration_frac=ib / (ms / iters / 1e3) / 1e9 / HBM_PEAK_GBS, grid_cell=ch["grid_cell"],
                choice=ch, loss_paths_start=stats0, loss_paths_end=stats1,
                loss_kernel_algorithmic_bytes=lb, extract=dict(wall_ms=extract_wall, **mesh_summary(mesh)),
                fp32_vs_fp64=gate, grad_clamped_targets=saturated)


def run_small(iters=200, K=16):
    """Plain single fit: Poisson r=0.025 seed 0, 8,192 star targets (sigma 1e-3), 200 iterations, clip with the
    star's occupancy; gradient gate at the fitted sites."""
    import vmhip
    from vmhip import workloads as wl
    r = 0.025
    smp = vmhip.Sampler(1.0, 1.0, r, 30, 0)
    grid, gw, gh, cs = smp.cells(), smp.grid_w, smp.grid_h, smp.cell_size
    tg = wl.star(8192, sigma=1e-3, seed=0)
    sites = wl.live_sites(grid)
    cell = 2.0 / math.sqrt(len(sites))
    opt = vmhip.Optimizer(1.0, 1.0, r, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / r, iterations=iters,
                          grid_cell=cell)
    opt.set_targets(tg)
    opt.set_generators(grid, gw, gh, cs)
    masks, tx, ty, vox = occupancy([wl.star_outline()])
    opt.set_occupancy(masks, tx, ty, vox)
    gate_start = single_gate(grid, gw, gh, cs, r, tg, K, grid_cell=cell)  # at the Poisson start
    t0 = time.perf_counter()
    gpu_ms = opt.optimize()
    mesh = opt.extract()
    wall = time.perf_counter() - t0
    fitted = opt.generators()
    out = dict(sites=len(sites), targets=len(tg), iterations=iters, gpu_ms=gpu_ms,
               gpu_site_iters_per_s=len(sites) * iters / (gpu_ms / 1e3), wall_s=wall,
               final_mean_loss=opt.evaluate_loss(), extract=mesh_summary(mesh),
               grad_clamped_targets=opt.grad_saturation())
    opt.close()
    out["fp32_vs_fp64_start"] = gate_start
    out["fp32_vs_fp64"] = single_gate(fitted, gw, gh, cs, r, tg, K, grid_cell=cell)  # at the fitted sites
    return out, (sites, tg, r, gw, gh, cs, K, iters, cell)


def run_precision(iters=5, K=16):
    """Jittered 256x256 lattice (+-1e-7) and 524,288 circle targets: the FP32 run (FP64 clip and loss sum)
    against the all-FP64 GPU run: gradient (with the FP32 run's choices), positions after `iters` steps,
    clipped-cell topology, and the circle's boundary.  (Both runs against the FP64 CPU oracle are tests.)"""
    import vmhip
    from vmhip import workloads as wl
    sites = wl.lattice(256, 1e-7, seed=3)
    tg = wl.circle(524288, radius=0.3, sigma=1e-4, seed=3)
    masks, tx, ty, vox = occupancy([wl.circle_outline(0.3)])
    runs = {}
    for prec in (0, 1):
        opt = vmhip.Optimizer(1.0, 1.0, 1 / 256, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 * 256,
                              iterations=iters, precision=prec)
        opt.set_targets(tg)
        opt.set_generators(sites, 256, 256, 1 / 256)
        opt.set_occupancy(masks, tx, ty, vox)
        lp = opt.loss_pass(len(tg))
        mesh0 = opt.extract()
        opt.step(iters)
        runs[prec] = (lp, mesh0, opt.generators().astype(np.float64), opt.grad_saturation())
        opt.close()
    lp64, m64, p64, _ = runs[1]
    lp32, mesh0, pos, saturated = runs[0]
    pe = np.abs(pos - p64)
    mism = 0
    for i in range(len(sites)):
        a = mesh0["cell_labels"][mesh0["cell_start"][i]:mesh0["cell_start"][i + 1]]
        b = m64["cell_labels"][m64["cell_start"][i]:m64["cell_start"][i + 1]]
        mism += int(not np.array_equal(a, b))
    gate = gradient_gate([sites], [tg], lp32, lp64)
    gate.update(pos_max_abs_err=float(pe.max()), pos_mean_abs_err=float(pe.mean()), topology_mismatches=mism)
    return dict(fp32_vs_fp64=gate, extract_fp32=mesh_summary(mesh0), extract_fp64=mesh_summary(m64),
                sites=len(sites), targets=len(tg), iterations=iters, grad_clamped_targets=saturated)


def _shot(args, timeout, full=False):
    import subprocess
    exe = os.path.join(ROOT, "tests", "cpp", "shot_demo")
    env = dict(os.environ, VM_EDIT_FULL="1" if full else "0")
    r = subprocess.run([exe, *args, "--json"], capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"shot_demo failed ({r.returncode}): {r.stderr[-500:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def run_local_edit(timeout=240):
    """The demo's scripted edit sequence, headless (tests/cpp/shot_demo): per-stamp latency of the local pipeline
    (canvas, upload, Poisson resample, 100 local iterations, extract) on a 1024^2 canvas with Poisson r = 24
    generators; and on a 4096^2 canvas with r = 1.63 (889^2 generator slots, about 263k live sites, the large
    instance's size), with local edits patched in place and, for comparison, with every edit re-laid out.
    Run as child processes."""
    if not os.path.exists(os.path.join(ROOT, "tests", "cpp", "shot_demo")):
        return {"skipped": "tests/cpp/shot_demo not built"}
    out = _shot(["24", "1024", "100"], timeout)
    big = _shot(["1.63", "4096", "100"], timeout)
    big_full = _shot(["1.63", "4096", "100"], timeout, full=True)
    keys = ("canvas_ms", "upload_ms", "poisson_ms", "optimize_ms", "extract_ms", "total_ms")
    out["large_canvas"] = dict(radius=1.63, raster=4096, generator_slots=big["generator_slots"],
                               live_sites=big["heart"]["live"], stamps=big["stamps"],
                               incremental={k: big[k]["mean"] for k in keys},
                               relayout_every_edit={k: big_full[k]["mean"] for k in keys},
                               watertight_after=big["after"]["watertight"])
    return out


def serving_summary(rec, st, rate, seconds, warm=2.0):
    """One window of the synthetic stream: throughput (site-iterations executed, and those of completed requests),
    latency percentiles of the requests that arrived after `warm` s and completed, and the backlog at the end."""
    done = rec["done_s"] >= 0
    adm = rec["admit_s"] >= 0
    sel = done & (rec["arrival_s"] >= warm)
    lat = (rec["done_s"] - rec["arrival_s"])[sel]
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else None  # noqa: E731
    wall = st["wall_s"]
    return dict(
        arrival_rate=rate, window_s=seconds, wall_s=wall, arrived=st["arrived"], admitted=st["admitted"],
        completed=st["completed"], backlog_at_end=st["arrived"] - st["admitted"],
        in_flight_at_end=st["admitted"] - st["completed"], oldest_wait_at_end_s=st["oldest_wait_s"],
        p50_latency_s=pct(lat, 50), p99_latency_s=pct(lat, 99), p50_queue_s=pct((rec["admit_s"] - rec["arrival_s"])[sel], 50),
        requests_per_s=st["completed"] / wall, admissions_per_s=st["admitted"] / wall,
        site_iters_per_s=st["site_iters_run"] / wall, completed_site_iters_per_s=st["site_iters"] / wall,
        gpu_busy_frac=st["gpu_busy_s"] / wall, mean_sites=float(rec["num_sites"][adm].mean()) if adm.any() else None,
        mean_iterations=float(rec["iterations"][done].mean()) if done.any() else None,
        max_resident=st["max_resident_seen"], peak_resident_GB=st["peak_bytes"] / 1e9, pool_GB=st["pool_bytes"] / 1e9,
        pool_grows=st["pool_grows"], upload_GB=st["upload_bytes"] / 1e9,
        host=dict(prep_workers=st["prep_workers"], prep_s=st["prep_busy_s"], prep_wait_s=st["prep_wait_s"],
                  sampler_s=st["sample_busy_s"], loop_s=st["loop_busy_s"]))


def run_serving(seconds, rate, cap, seed=4, budget=200e9, max_iters=300, tol=1e-6):
    """The serving stream: Poisson arrivals at `rate`/s for `seconds`, each request generated inside the library when
    it is about to be admitted (site count log-uniform in [512, 131072], Poisson-disk sites, 8 targets per site on a
    random polygon, sigma 5e-4), run to |dL/L| < tol or max_iters; 200 GB budget, at most `cap` resident sites.  The
    window ends at `seconds` (no drain)."""
    import vmhip
    stream = vmhip.serve_stream_desc(rate=rate, duration_s=seconds, seed=seed, drain=False)
    rec, st = vmhip.serve_synthetic(stream, hbm_budget_bytes=budget, max_iterations=max_iters, tol=tol,
                                    poll_every=10, max_resident_sites=cap)
    return serving_summary(rec, st, rate, seconds), rec


def serving_sustained(rec, seconds, oldest_wait, warm=2.0, settle=4.0, max_wait=0.25):
    """Bounded p99: every request that arrived before the window's last `settle` s completed, the p99 latency of
    those arriving in the later half of [warm, seconds - settle] is within 1.5x of the earlier half's, and no
    request had waited more than `max_wait` s for admission when the window ended (a queue that grows slowly
    shows there first)."""
    end = seconds - settle
    mid = 0.5 * (warm + end)
    early = (rec["arrival_s"] >= warm) & (rec["arrival_s"] < mid)
    late = (rec["arrival_s"] >= mid) & (rec["arrival_s"] < end)
    if not (np.all(rec["done_s"][early | late] >= 0) and early.any() and late.any()):
        return False, None, None
    lat = rec["done_s"] - rec["arrival_s"]
    p_early, p_late = float(np.percentile(lat[early], 99)), float(np.percentile(lat[late], 99))
    return p_late <= 1.5 * p_early and oldest_wait <= max_wait, p_early, p_late


def run_serving_sweep(rates, seconds, cap):
    out = []
    for r in rates:
        summ, rec = run_serving(seconds, r, cap)
        ok, pe, pl = serving_sustained(rec, seconds, summ["oldest_wait_at_end_s"])
        summ.update(sustained=ok, p99_early_s=pe, p99_late_s=pl)
        out.append(summ)
    return out


def serving_plain(count=16, iters=20, seed=4):
    """The plain one-fit-at-a-time path on requests of the serving stream (regenerated from it): site-iters/s."""
    import vmhip
    stream = vmhip.serve_stream_desc(rate=2000.0, duration_s=60.0, seed=seed)
    insts = []
    for i in range(count):
        req, _ = vmhip.synth_request(stream, i)
        insts.append((req["sites"], req["targets"]))
    rate, _ = run_plain(insts, iters)
    return rate, float(np.mean([len(s) for s, _ in insts]))


# ---------------------------------------------------------------------------
# CPU baseline (the oracle restatement on the host's cores)
# ---------------------------------------------------------------------------
def cpu_fit_rate(cases, iters, threads=None):
    import oracle as O
    if threads:
        O.set_threads(threads)
    sc = O.make_sched(1e-3, 1.0, 1.0)
    t0 = time.perf_counter()
    si = 0
    for s, t, K in cases:
        O.fit(s, t, K, 2.0 / math.sqrt(len(s)), sc, iters)
        si += len(s) * iters
    dt = time.perf_counter() - t0
    return si / dt, dt


def cpu_baselines(insts, ks, small_case, large_case, prec_case, iters=2):
    """Batch: every instance of the 512-fit workload at `iters` iterations on all host cores (the headline baseline),
    and every 16th instance on one thread; the small fit in full (200 iterations and the clip); the large instance and
    the precision stress at 3 iterations each."""
    import oracle as O
    threads = O.num_threads()
    cases = [(s, t, k) for (s, t), k in zip(insts, ks)]
    rate, dt = cpu_fit_rate(cases, iters)
    rate1, dt1 = cpu_fit_rate(cases[::16], 1, threads=1)
    O.set_threads(threads)
    sites, tg, r, gw, gh, cs, K, sit, cell = small_case
    lr = float(np.float32(np.float32(1e-3 / r) * np.float32(r)))
    m = np.float32(2) * np.float32(r)
    box = (float(-m), float(-m), float(np.float32(gw) * np.float32(cs) + m), float(np.float32(gh) * np.float32(cs) + m))
    t0 = time.perf_counter()
    ref, _, _, _, _ = O.fit(sites, tg, K, cell, O.make_sched(lr, 1.0, 1.0), sit, box=box)
    O.clip(ref, 1.0, 1.0, cell)
    small_s = time.perf_counter() - t0
    lrate, ldt = cpu_fit_rate([large_case], 3)
    prate, pdt = cpu_fit_rate([prec_case], 3)
    return dict(value=rate, unit="site-iters/s", cores=threads, kind="port",
                sample=f"all {len(insts)} instances of the batch ({sum(len(s) for s, _ in insts)} sites), {iters} "
                       f"iterations each, {dt:.1f} s, oracle restatement (OpenMP {threads} threads, {cpu_model()})",
                one_thread=dict(value=rate1, sample=f"every 16th instance ({len(cases[::16])}), 1 iteration, "
                                                     f"{dt1:.1f} s"),
                small_fit=dict(s=small_s, site_iters_per_s=len(sites) * sit / small_s,
                               sample=f"{len(sites)} sites, {sit} iterations and the clip"),
                large_instance=dict(site_iters_per_s=lrate, s=ldt, sample="3 iterations"),
                precision_stress=dict(site_iters_per_s=prate, s=pdt, sample="3 iterations"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--instances", type=int, default=512)
    ap.add_argument("--plain-iters", type=int, default=20)
    ap.add_argument("--plain-instances", type=int, default=0, help="0: all instances")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="arrival window of the serving run (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-cap", type=float, default=8e6, help="resident-site cap of the serving loop")
    ap.add_argument("--serve-sweep", default="500,700,800,900,950,1000", help="arrival rates of the sustainable-rate sweep")
    ap.add_argument("--serve-sweep-seconds", type=float, default=14.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the large/small/precision/local-edit sections")
    ap.add_argument("--only", choices=["large", "small", "precision", "local_edit", "batch", "serving"], default=None,
                    help="run one section alone and print its JSON (profiling)")
    ap.add_argument("--large-iters", type=int, default=500)
    ap.add_argument("--no-gate", action="store_true", help="skip the FP64 gate and the meshes (profiling runs)")
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
    if args.only and args.only != "batch":
        fn = {"large": lambda: run_large(iters=args.large_iters), "small": lambda: run_small()[0],
              "precision": run_precision, "local_edit": run_local_edit, "serving": lambda: serving_section(args)}
        print(json.dumps(fn[args.only](), default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
        return
    log(f"generating {args.instances} instances ...")
    t0 = time.time()
    insts, seeds, poisson_ms = make_batched(args.instances, args.seed)
    S = sum(len(s) for s, _ in insts)
    T = sum(len(t) for _, t in insts)
    log(f"  {S} sites, {T} targets in {time.time() - t0:.1f} s (Poisson {poisson_ms:.1f} ms GPU)")

    log(f"batched (tuned k): warmup {args.warmup}, timed {args.steps} iterations ...")
    b, ms = run_batched(insts, args.steps, args.warmup, loss_k=0)
    value = S * args.steps / (ms / 1e3)
    ms_per_step = ms / args.steps
    phase = b.profile(5)
    paths = b.loss_stats()
    C = batch_cells(b, len(insts))
    ks = [b.choice(i)["loss_k"] for i in range(len(insts))]
    tpls = [b.choice(i)["targets_per_lane"] for i in range(len(insts))]
    log(f"  {ms_per_step:.3f} ms/iteration, {value:.3e} site-iters/s; phases ms {np.round(phase, 4).tolist()}")
    gate, meshes = None, None
    if not args.no_gate:
        log("  gradient gate (FP32 vs FP64 with the FP32 choices) ...")
        gate = batch_gate(b, insts)
        gate["grad_clamped_targets"] = b.saturation()
        log(f"  {gate}")
        meshes = batch_meshes(b, insts, seeds)
    b.close()
    if args.only == "batch":
        print(json.dumps(dict(value=value, ms_per_step=ms_per_step, phases=phase_report(phase, S, T, C), gate=gate,
                              paths=paths), default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)
        return

    log("batched with k fixed at 16 ...")
    b16, ms16 = run_batched(insts, args.steps, args.warmup, loss_k=16)
    phase16 = b16.profile(5)
    b16.close()

    npl = args.plain_instances or len(insts)
    log(f"plain: {npl} fits one at a time, {args.plain_iters} iterations each ...")
    plain_rate, plain_ms = run_plain(insts[:npl], args.plain_iters)
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
                               "random polygon (sigma 5e-4), lr 1e-3, k/cell/chunk shape tuned per size class, one "
                               "persistent set of launches per iteration (HIP graph)",
                   "instances": len(insts), "sites": S, "targets": T, "hash_cells": C,
                   "k_histogram": {str(k): ks.count(k) for k in sorted(set(ks))},
                   "targets_per_lane_histogram": {str(t): tpls.count(t) for t in sorted(set(tpls))}},
        "grad_gate": gate,
        "plain_site_iters_per_s": plain_rate,
        "batched_over_plain": value / plain_rate,
        "k16": {"site_iters_per_s": S * args.steps / (ms16 / 1e3), "ms_per_step": ms16 / args.steps,
                "kernels": phase_report(phase16, S, T, C)},
        "loss_paths": paths,
        "kernels": phase_report(phase, S, T, C),
        "iteration_roofline": {"bytes": iter_bytes(S, T, C),
                               "achieved_GBs": iter_bytes(S, T, C) / (ms_per_step / 1e3) / 1e9},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "lossKernel (batch)",
                     "algorithmic_bytes": lb, "kernel_ms": loss_ms},
        "batch_meshes": meshes,
    }
    traffic = measured_traffic()
    if traffic and args.instances == 512 and args.seed == 2:
        out["roofline"]["traffic"] = traffic
    small_case = large_case = prec_case = None
    if not args.no_extra:
        log("config 1: large single instance ...")
        out["large_instance"] = run_large(iters=args.large_iters)
        li = out["large_instance"]
        log(f"  {li['ms_per_iter']:.4f} ms/iter, {li['site_iters_per_s']:.3e} site-iters/s; gate {li['fp32_vs_fp64']}")
        log("config 0: small plain fit ...")
        out["small"], small_case = run_small()
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
        from vmhip import workloads as wl
        large_case = (wl.live_sites(vmhip.poisson_batch([1.0], [1.0], [wl.poisson_radius_for(262144)], [1])[0][0]),
                      wl.ellipses(4194304, count=64, sigma=2e-4, seed=1), 16)
        prec_case = (wl.lattice(256, 1e-7, seed=3), wl.circle(524288, radius=0.3, sigma=1e-4, seed=3), 16)
    if args.serve_seconds > 0:
        out["serving"] = serving_section(args)
    if not args.no_cpu and small_case is not None:
        log("cpu baselines (oracle) ...")
        out["cpu_baseline"] = cpu_baselines(insts, ks, small_case, large_case, prec_case)
        log(f"  cpu {out['cpu_baseline']['value']:.3e} site-iters/s")
    print(json.dumps(out, default=lambda o: o.item() if hasattr(o, "item") else str(o)), flush=True)


def serving_section(args):
    """The serving stream at --serve-rate for --serve-seconds, the sustainable-rate sweep, the plain one-at-a-time
    path on the same law, and the demand-vs-capacity arithmetic."""
    log(f"serving: Poisson arrivals {args.serve_rate}/s for {args.serve_seconds} s, cap {args.serve_cap:.0f} sites ...")
    sv, _ = run_serving(args.serve_seconds, args.serve_rate, args.serve_cap)
    log(f"  {sv['arrived']} arrived, {sv['admitted']} admitted, {sv['completed']} completed; "
        f"{sv['site_iters_per_s']:.3e} site-iters/s, p50 {sv['p50_latency_s']:.3f} s, p99 {sv['p99_latency_s']:.3f} s, "
        f"backlog {sv['backlog_at_end']}")
    rates = [float(r) for r in args.serve_sweep.split(",") if r]
    log(f"serving sweep: {rates} for {args.serve_sweep_seconds} s each ...")
    sweep = run_serving_sweep(rates, args.serve_sweep_seconds, args.serve_cap)
    for r in sweep:
        log(f"  {r['arrival_rate']:.0f}/s: p50 {r['p50_latency_s']}, p99 {r['p99_latency_s']}, sustained {r['sustained']}, "
            f"backlog {r['backlog_at_end']}, {r['site_iters_per_s']:.3e} site-iters/s")
    ok = [r["arrival_rate"] for r in sweep if r["sustained"]]
    log("serving: plain one-fit-at-a-time path on 16 requests of the stream ...")
    plain, plain_mean_sites = serving_plain()
    es, eit = sv["mean_sites"], sv["mean_iterations"]
    per_req = es * eit
    return dict(
        config=dict(arrival_rate=args.serve_rate, window_s=args.serve_seconds, sites="log-uniform [512, 131072]",
                    targets_per_site=8, sigma=5e-4, stop="|dL/L| < 1e-6 or 300 iterations", hbm_budget_GB=200,
                    max_resident_sites=args.serve_cap, k=16, generation="lazy, inside the library, at admission"),
        window=sv, sweep=sweep, highest_sustained_rate=max(ok) if ok else None,
        plain_one_at_a_time=dict(site_iters_per_s=plain, sample="16 requests of the stream, 20 iterations each",
                                 mean_sites=plain_mean_sites),
        vs_plain_one_at_a_time=sv["site_iters_per_s"] / plain,
        demand=dict(mean_sites=es, mean_iterations=eit, site_iters_per_request=per_req,
                    demand_site_iters_per_s=args.serve_rate * per_req,
                    capacity_site_iters_per_s=sv["site_iters_per_s"],
                    capacity_requests_per_s=sv["site_iters_per_s"] / per_req,
                    plain_capacity_requests_per_s=plain / per_req,
                    load_factor=args.serve_rate * per_req / sv["site_iters_per_s"]))


def measured_traffic():
    """HBM bytes per batch loss-kernel launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    (profiles/loss_kernel_traffic.json, calibrated on a 256 MiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "loss_kernel_traffic.json")
    try:
        with open(path) as f:
            return float(json.load(f)["batch512"]["traffic_bytes"])
    except (OSError, KeyError, ValueError):
        return None


if __name__ == "__main__":
    main()
