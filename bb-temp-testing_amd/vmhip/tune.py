"""Tuner for the loss kernel, per size class of fits.

For each size class (fits of 2^c .. 2^(c+1) - 1 sites), a batch of fits of that class (Poisson-disk sites, 8
targets per site on random polygons) is run on the GPU, and the choices with the highest site-iterations/s are
written to csrc/vm_tune_table.h, which the library compiles in.  Three stages per class:

  1. hash-grid cell scale at k = 16, chunk extent 2.2 and 3 targets per lane;
  2. targets per lane (3, 4) x chunk extent at the chosen cell scale;
  3. k for fits that leave it to the tuner: among 8, 12 and 16 the ones whose per-target loss terms (nearest
     site, bisector partner, term) equal those of k = 32 on every target of the class's check fits, both at the
     Poisson start and after 50 iterations (so the choice cannot change a result on them), 16 always; the
     smallest of them (the least work per target), unless a larger one runs faster.

Classes above the largest measured one reuse its row.  Run on a GPU box, then rebuild:

    python -m vmhip.tune --out bb-temp-testing_amd/csrc/vm_tune_table.h
"""
import argparse
import math
import sys
import time

import numpy as np

import vmhip
from vmhip import workloads as wl

CELL_SCALES = (0.7, 0.8, 0.9, 1.0, 1.15, 1.3)
CHUNK_EXTENTS = (1.6, 2.2, 3.0, 4.0)
# Two targets per lane only pays in a batch of nothing but tiny fits (more, smaller chunks); in mixed batches the
# larger fits' chunks dominate, so the sweep keeps to 3 and 4.
TARGETS_PER_LANE = (3, 4)
K_CANDIDATES = (8, 12, 16)
K_REFERENCE = 32
MIN_GAIN = 0.015
LOG2_MIN, LOG2_MAX = 8, 22


def base_cell(K):
    """Base hash-grid cell in mean spacings (the library's formula; the table scales it)."""
    return max(1.5, 0.5 * (math.sqrt(K / math.pi) + 0.8))


def make_class(c, budget_sites, max_fits, seed):
    """Fits of 2^c .. 2^(c+1) - 1 sites (log-uniform), about budget_sites sites in total, with their sampler grids."""
    lo, hi = 2 ** c, 2 ** (c + 1) - 1
    n = int(min(max_fits, max(1, budget_sites // int(1.5 * lo))))
    sizes = wl.log_uniform_sizes(n, lo, hi, seed)
    radii = np.array([wl.poisson_radius_for(int(s)) for s in sizes], np.float32)
    seeds = (7 * (1 << 20) + 4096 * c + np.arange(n)).astype(np.uint32)
    grids, dims, _ = vmhip.poisson_batch(np.ones(n), np.ones(n), radii, seeds)
    fits = []
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        fits.append(dict(sites=sites, targets=wl.polygon(8 * len(sites), sigma=5e-4, seed=int(seeds[i])),
                         grid=g, dims=dims[i], radius=float(radii[i])))
    return fits


def measure(fits, configs, iters, rounds=5):
    """Site-iterations/s of the class batch under each configuration (K, cell scale, extent, targets per lane).
    Every configuration's batch is built first; then `rounds` rounds time `iters` iterations (graph replays of
    the whole batch) of each batch in turn, so drift of the clocks between configurations cancels, and each
    configuration takes the median of its rounds."""
    batches = []
    for K, scale, extent, tpl in configs:
        b = vmhip.Batch()
        for f in fits:
            a = math.sqrt(1.0 / len(f["sites"]))
            b.add(1.0, 1.0, f["sites"], f["targets"], 1e-3, loss_k=K, grid_cell=scale * base_cell(K) * a,
                  stage_rings=2, chunk_extent=extent, targets_per_lane=tpl)
        b.commit()
        b.step(3)
        batches.append(b)
    times = [[] for _ in batches]
    for _ in range(rounds):
        for i, b in enumerate(batches):
            times[i].append(b.step(iters))
    for b in batches:
        b.close()
    sites = sum(len(f["sites"]) for f in fits)
    return [sites * iters / (sorted(t)[len(t) // 2] / 1e3) for t in times]


def terms(sites_slots, sw, sh, cell, targets, K, radius):
    """Per-target (nearest, partner, term) of one loss pass at k = K on the optimizer's own path."""
    opt = vmhip.Optimizer(1.0, 1.0, radius, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / radius)
    opt.set_targets(targets)
    opt.set_generators(sites_slots, sw, sh, cell)
    d = opt.loss_pass(len(targets))
    opt.close()
    return d["nearest"], d["partner"], d["loss_terms"]


def k_agreement(fits, Ks, iters=50):
    """For each K: True when k = K gives the same per-target terms as k = 32 on every target of the check fits,
    at the start and after `iters` iterations of a k = 16 batch."""
    b = vmhip.Batch()
    for f in fits:
        b.add(1.0, 1.0, f["sites"], f["targets"], 1e-3, loss_k=16)
    b.commit()
    b.step(iters)
    ok = {K: True for K in Ks}
    for i, f in enumerate(fits):
        # Live sites as a row of slots (a sampler grid of the largest classes has more slots than an optimizer
        # takes; empty slots play no part in the terms).
        start, moved = f["sites"], b.sites(i)
        for slots, sw, sh, cell in ((start, len(start), 1, 1.0 / len(start)), (moved, len(moved), 1, 1.0 / len(moved))):
            ref = terms(slots, sw, sh, cell, f["targets"], K_REFERENCE, f["radius"])
            for K in Ks:
                if ok[K]:
                    got = terms(slots, sw, sh, cell, f["targets"], K, f["radius"])
                    ok[K] = all(np.array_equal(x, y) for x, y in zip(ref, got))
    b.close()
    return ok


def write_table(path, rows):
    lines = ["// vm_tune_table.h -- per-size-class choices of the tuner (written by",
             "// bb-temp-testing_amd/vmhip/tune.py from a measured sweep; edit by re-running it).",
             "// Row i covers fits of 2^(kTuneLog2Min + i) .. 2^(kTuneLog2Min + i + 1) - 1 sites.",
             "#pragma once", "", "namespace vmrt {", "",
             "struct TuneRow {",
             "    float cellScale;   // multiple of the base hash-grid cell max(1.5, (sqrt(K/pi) + 0.8) / 2) spacings",
             "    float chunkExtent; // target chunk diagonal, in mean site spacings",
             "    int   tpl;         // loss kernel targets per lane",
             "    int   k;           // loss neighbours when a fit leaves k to the tuner", "};", "",
             f"constexpr int     kTuneLog2Min = {LOG2_MIN};", f"constexpr int     kTuneLog2Max = {LOG2_MAX};",
             "constexpr TuneRow kTuneTable[kTuneLog2Max - kTuneLog2Min + 1] = {"]
    for c in range(LOG2_MIN, LOG2_MAX + 1):
        r = rows[c]
        note = (f"{r['rate']:.3e} site-iters/s at k=16, {r['rate_k']:.3e} at k={r['k']}" if r.get("rate")
                else "largest measured class")
        lines.append(f"    {{{r['scale']:.2f}f, {r['extent']:.2f}f, {r['tpl']}, {r['k']}}}, // 2^{c}: {note}")
    lines += ["};", "", "} // namespace vmrt", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def tune_class(c, args):
    """Coordinate search from the default (cell x1.0, extent 2.2, 3 targets per lane, k = 16); a candidate replaces
    the current choice only when it is faster by more than MIN_GAIN (run-to-run noise is about 1 %).  The
    candidates of each stage are timed in interleaved rounds."""
    t0 = time.time()
    budget = max(args.budget, int(1.5 * 2 ** c))
    fits = make_class(c, budget, args.max_fits, seed=c)
    say = lambda m: print(f"class 2^{c} ({len(fits)} fits): {m}", flush=True)  # noqa: E731
    rates = measure(fits, [(16, sc, 2.2, 3) for sc in CELL_SCALES], args.iters)
    rate1 = dict(zip(CELL_SCALES, rates))
    best = dict(scale=1.0, extent=2.2, tpl=3, rate=rate1[1.0])
    for sc in CELL_SCALES:
        say(f"cell x{sc:.2f}, extent 2.2, tpl 3: {rate1[sc]:.3e}")
        if rate1[sc] > best["rate"] * (1 + MIN_GAIN):
            best.update(scale=sc, rate=rate1[sc])
    shapes = [(tpl, e) for tpl in TARGETS_PER_LANE for e in CHUNK_EXTENTS]
    rates = measure(fits, [(16, best["scale"], e, tpl) for tpl, e in shapes], args.iters)
    rate2 = dict(zip(shapes, rates))
    best.update(rate=rate2[(3, 2.2)])
    for (tpl, e) in shapes:
        say(f"cell x{best['scale']:.2f}, extent {e:.1f}, tpl {tpl}: {rate2[(tpl, e)]:.3e}")
        if rate2[(tpl, e)] > best["rate"] * (1 + MIN_GAIN):
            best.update(extent=e, tpl=tpl, rate=rate2[(tpl, e)])
    # Check fits: a few of the class (one for classes of a million sites and more: each pass holds its k-NN lists).
    check = fits[:min(len(fits), args.check_fits if c < 19 else 1)]
    agree = k_agreement(check, [K for K in K_CANDIDATES if K != 16])
    agree[16] = True
    ks = [K for K in K_CANDIDATES if agree[K]]
    rates = dict(zip(ks, measure(fits, [(K, best["scale"], best["extent"], best["tpl"]) for K in ks], args.iters)))
    for K in K_CANDIDATES:
        say(f"k {K}: {rates.get(K, float('nan')):.3e}, terms equal to k={K_REFERENCE}'s: {agree[K]}")
    # The smallest agreeing k; a larger one replaces it only when faster by more than MIN_GAIN.
    ks = sorted(rates)
    best.update(k=ks[0], rate_k=rates[ks[0]])
    for K in ks[1:]:
        if rates[K] > best["rate_k"] * (1 + MIN_GAIN):
            best.update(k=K, rate_k=rates[K])
    say(f"best cell x{best['scale']:.2f}, extent {best['extent']:.1f}, tpl {best['tpl']}, k {best['k']} "
        f"({time.time() - t0:.1f} s)")
    return best


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--classes", default="8-20", help="measured size classes (log2 of the site count)")
    ap.add_argument("--budget", type=int, default=2_000_000, help="sites per class (enough to fill the GPU, as a mixed batch does)")
    ap.add_argument("--max-fits", type=int, default=2048)
    ap.add_argument("--check-fits", type=int, default=3, help="fits per class for the k agreement check")
    ap.add_argument("--iters", type=int, default=100, help="iterations per timed run (about 30 ms of a class batch)")
    args = ap.parse_args()
    c0, c1 = (int(x) for x in args.classes.split("-"))
    rows = {c: tune_class(c, args) for c in range(c0, c1 + 1)}
    for c in range(LOG2_MIN, LOG2_MAX + 1):
        if c not in rows:
            near = min(rows, key=lambda m: abs(m - c))
            rows[c] = {k: v for k, v in rows[near].items() if k not in ("rate", "rate_k")}
    write_table(args.out, rows)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    sys.exit(main())
