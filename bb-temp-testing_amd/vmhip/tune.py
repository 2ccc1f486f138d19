"""Tuner sweep for the loss kernel's geometry, per size class of fits.

For each size class (fits of 2^c .. 2^(c+1) - 1 sites), a batch of fits of that class (Poisson-disk sites, 8 targets
per site on random polygons, k=16) is run on the GPU for every candidate pair (hash-grid cell scale, target chunk
extent); the pair with the highest site-iterations/s is written to csrc/vm_tune_table.h, which the library compiles
in.  Classes above the largest measured one reuse its row.  Run on a GPU box, then rebuild:

    python -m vmhip.tune --out bb-temp-testing_amd/csrc/vm_tune_table.h
"""
import argparse
import math
import sys
import time

import numpy as np

import vmhip
from vmhip import workloads as wl

CELL_SCALES = (0.9, 1.0, 1.15)
CHUNK_EXTENTS = (1.0, 1.6, 2.2, 3.0)
LOG2_MIN, LOG2_MAX = 8, 22


def base_cell(K):
    """Base hash-grid cell in mean spacings (the library's formula; the table scales it)."""
    return max(1.5, 0.5 * (math.sqrt(K / math.pi) + 0.8))


def make_class(c, budget_sites, max_fits, seed):
    """Fits of 2^c .. 2^(c+1) - 1 sites (log-uniform), about budget_sites sites in total."""
    lo, hi = 2 ** c, 2 ** (c + 1) - 1
    n = int(min(max_fits, max(1, budget_sites // int(1.5 * lo))))
    sizes = wl.log_uniform_sizes(n, lo, hi, seed)
    radii = np.array([wl.poisson_radius_for(int(s)) for s in sizes], np.float32)
    seeds = (7 * (1 << 20) + 4096 * c + np.arange(n)).astype(np.uint32)
    grids, _, _ = vmhip.poisson_batch(np.ones(n), np.ones(n), radii, seeds)
    fits = []
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        fits.append((sites, wl.polygon(8 * len(sites), sigma=5e-4, seed=int(seeds[i]))))
    return fits


def measure(fits, K, scale, extent, iters):
    b = vmhip.Batch()
    for sites, tg in fits:
        a = math.sqrt(1.0 / len(sites))
        b.add(1.0, 1.0, sites, tg, 1e-3, loss_k=K, grid_cell=scale * base_cell(K) * a, stage_rings=2,
              chunk_extent=extent)
    b.commit()
    b.step(3)
    ms = b.step(iters)
    b.close()
    return sum(len(s) for s, _ in fits) * iters / (ms / 1e3)


def write_table(path, rows):
    lines = ["// vm_tune_table.h -- per-size-class choices of the tuner (written by",
             "// bb-temp-testing_amd/vmhip/tune.py from a measured sweep; edit by re-running it).",
             "// Row i covers fits of 2^(kTuneLog2Min + i) .. 2^(kTuneLog2Min + i + 1) - 1 sites.",
             "#pragma once", "", "namespace vmrt {", "",
             "struct TuneRow {",
             "    float cellScale;   // multiple of the base hash-grid cell max(1.5, (sqrt(K/pi) + 0.8) / 2) spacings",
             "    float chunkExtent; // target chunk diagonal, in mean site spacings", "};", "",
             f"constexpr int     kTuneLog2Min = {LOG2_MIN};", f"constexpr int     kTuneLog2Max = {LOG2_MAX};",
             "constexpr TuneRow kTuneTable[kTuneLog2Max - kTuneLog2Min + 1] = {"]
    for c in range(LOG2_MIN, LOG2_MAX + 1):
        s, e, rate = rows[c]
        note = f"measured {rate:.3e} site-iters/s" if rate else "largest measured class"
        lines.append(f"    {{{s:.2f}f, {e:.2f}f}}, // 2^{c}: {note}")
    lines += ["};", "", "} // namespace vmrt", ""]
    with open(path, "w") as f:
        f.write("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--classes", default="8-18", help="measured size classes (log2 of the site count)")
    ap.add_argument("--budget", type=int, default=1_000_000, help="sites per class")
    ap.add_argument("--max-fits", type=int, default=512)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--k", type=int, default=16)
    args = ap.parse_args()
    c0, c1 = (int(x) for x in args.classes.split("-"))
    rows = {}
    for c in range(c0, c1 + 1):
        t0 = time.time()
        fits = make_class(c, args.budget, args.max_fits, seed=c)
        best = None
        for s in CELL_SCALES:
            for e in CHUNK_EXTENTS:
                rate = measure(fits, args.k, s, e, args.iters)
                print(f"class 2^{c} ({len(fits)} fits): cell x{s:.2f}, extent {e:.1f}: {rate:.3e} site-iters/s",
                      flush=True)
                if best is None or rate > best[2]:
                    best = (s, e, rate)
        rows[c] = best
        print(f"class 2^{c}: best cell x{best[0]:.2f}, extent {best[1]:.1f} ({time.time() - t0:.1f} s)", flush=True)
    for c in range(LOG2_MIN, LOG2_MAX + 1):
        if c not in rows:
            near = min(rows, key=lambda m: abs(m - c))
            rows[c] = (rows[near][0], rows[near][1], None)
    write_table(args.out, rows)
    print(f"wrote {args.out}")


if __name__ == "__main__":
    sys.exit(main())
