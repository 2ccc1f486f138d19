"""Seeded synthetic workloads for the VoroMesh fits (all in the unit square).

Target clouds are points spread uniformly by arc length over closed polylines
with isotropic Gaussian jitter:
  * ``star``      -- 5-point star, centre (0.5, 0.5), outer radius 0.40, inner 0.18
  * ``ellipses``  -- non-overlapping random ellipses
  * ``polygon``   -- random simple star-shaped polygon (sorted angles, random radii)
  * ``circle``    -- circle of radius 0.3
Sites come from the Poisson-disk sampler (``poisson_radius_for`` picks the
radius that yields about the requested count) or from a jittered lattice.
"""
from __future__ import annotations

import math

import numpy as np

# Sites per unit area times r^2 for the 30-trial dart thrower in the unit
# square: 0.684 at r = 0.025, 0.676 at 0.01, 0.670 at 0.005, tending to ~0.668
# for the large instances (measured on the sampler, seeds 0-3).
POISSON_YIELD = 0.668


def poisson_radius_for(num_sites: int, area: float = 1.0) -> float:
    return math.sqrt(POISSON_YIELD * area / num_sites)


def _on_polyline(verts: np.ndarray, n: int, sigma: float, rng: np.random.Generator) -> np.ndarray:
    v = np.asarray(verts, np.float64)
    seg = np.roll(v, -1, axis=0) - v
    ln = np.hypot(seg[:, 0], seg[:, 1])
    cum = np.concatenate([[0.0], np.cumsum(ln)])
    u = rng.random(n) * cum[-1]
    k = np.clip(np.searchsorted(cum, u, side="right") - 1, 0, len(v) - 1)
    t = (u - cum[k]) / np.where(ln[k] > 0, ln[k], 1.0)
    p = v[k] + t[:, None] * seg[k]
    p += rng.normal(0.0, sigma, size=p.shape)
    return p.astype(np.float32)


def star_outline(center=(0.5, 0.5), r_out=0.40, r_in=0.18):
    """The 5-point star's 10 vertices (outer vertex first, at angle pi/2)."""
    ang = math.pi / 2 + np.arange(10) * math.pi / 5
    rad = np.where(np.arange(10) % 2 == 0, r_out, r_in)
    return np.stack([center[0] + rad * np.cos(ang), center[1] + rad * np.sin(ang)], axis=1)


def star(n=8192, sigma=1e-3, seed=0, center=(0.5, 0.5), r_out=0.40, r_in=0.18):
    rng = np.random.default_rng([0x5747, seed])
    return _on_polyline(star_outline(center, r_out, r_in), n, sigma, rng)


def circle_outline(radius=0.3, center=(0.5, 0.5), segments=4096):
    a = np.arange(segments) * (2 * math.pi / segments)
    return np.stack([center[0] + radius * np.cos(a), center[1] + radius * np.sin(a)], axis=1)


def circle(n, radius=0.3, sigma=1e-4, seed=3, center=(0.5, 0.5), segments=4096):
    rng = np.random.default_rng([0xC1C, seed])
    return _on_polyline(circle_outline(radius, center, segments), n, sigma, rng)


def ellipse_outlines(count=64, seed=1, segments=4096):
    """The `count` ellipse polylines of `ellipses` (same placement draws)."""
    return _ellipse_shapes(count, np.random.default_rng([0xE11, seed]), segments)[0]


def _ellipse_shapes(count, rng, segments):
    shapes = []
    tries = 0
    while len(shapes) < count:
        tries += 1
        if tries > 1_000_000:
            raise RuntimeError("could not place the ellipses")
        c = rng.uniform(0.1, 0.9, 2)
        a, b = rng.uniform(0.01, 0.05, 2)
        th = rng.uniform(0.0, math.pi)
        rb = max(a, b)
        if any(math.hypot(*(c - s[0])) < rb + s[1] for s in shapes):
            continue
        shapes.append((c, rb, a, b, th))
    polys, perims = [], []
    t = np.arange(segments) * (2 * math.pi / segments)
    for c, _, a, b, th in shapes:
        x, y = a * np.cos(t), b * np.sin(t)
        ct, st = math.cos(th), math.sin(th)
        v = np.stack([c[0] + ct * x - st * y, c[1] + st * x + ct * y], axis=1)
        polys.append(v)
        perims.append(np.hypot(*(np.roll(v, -1, 0) - v).T).sum())
    return polys, np.asarray(perims)


def ellipses(n, count=64, sigma=2e-4, seed=1, segments=4096):
    """`count` non-overlapping ellipses (bounding circles disjoint), points split by perimeter."""
    rng = np.random.default_rng([0xE11, seed])
    polys, perims = _ellipse_shapes(count, rng, segments)
    counts = np.floor(n * perims / perims.sum()).astype(np.int64)
    counts[: n - counts.sum()] += 1
    return np.concatenate([_on_polyline(v, int(k), sigma, rng) for v, k in zip(polys, counts)])


def _polygon_verts(rng, center):
    m = int(rng.integers(5, 41))
    ang = np.sort(rng.uniform(0.0, 2 * math.pi, m))
    rad = rng.uniform(0.15, 0.45, m)
    return np.stack([center[0] + rad * np.cos(ang), center[1] + rad * np.sin(ang)], axis=1)


def polygon_outline(seed=2, center=(0.5, 0.5)):
    """The vertices of `polygon`'s random simple polygon for the same seed."""
    return _polygon_verts(np.random.default_rng([0xB01, seed]), center)


def polygon(n, sigma=5e-4, seed=2, center=(0.5, 0.5)):
    """Random simple polygon: 5-40 vertices at sorted U[0, 2pi) angles, radii U[0.15, 0.45]."""
    rng = np.random.default_rng([0xB01, seed])
    return _on_polyline(_polygon_verts(rng, center), n, sigma, rng)


def lattice(side=256, jitter=1e-7, seed=3):
    rng = np.random.default_rng([0x1A7, seed])
    i, j = np.meshgrid(np.arange(side), np.arange(side), indexing="xy")
    p = np.stack([(i.ravel() + 0.5) / side, (j.ravel() + 0.5) / side], axis=1)
    p += rng.uniform(-jitter, jitter, size=p.shape)
    return p.astype(np.float32)


def log_uniform_sizes(count, lo, hi, seed):
    rng = np.random.default_rng([0x512, seed])
    return np.round(np.exp(rng.uniform(math.log(lo), math.log(hi), count))).astype(np.int64)


def batch_plan(n=512, seed=2, lo=256, hi=65536):
    """The mixed batch's instances: Poisson radius (for a site count log-uniform in [lo, hi]) and sampler seed
    of each (seed 2 * 2^20 + i for instance i)."""
    sizes = log_uniform_sizes(n, lo, hi, seed)
    radii = np.array([poisson_radius_for(int(s)) for s in sizes], np.float32)
    seeds = (2 * (1 << 20) + np.arange(n)).astype(np.uint32)
    return radii, seeds


def batch_targets(sites, seed, per_site=8, sigma=5e-4):
    """Targets of one batch instance: `per_site` per site on its random polygon (seeded by the sampler seed)."""
    return polygon(per_site * len(sites), sigma=sigma, seed=int(seed))


def poisson_arrivals(rate, duration, seed=4):
    rng = np.random.default_rng([0xA77, seed])
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return np.asarray(out)
        out.append(t)


def live_sites(grid: np.ndarray) -> np.ndarray:
    g = np.asarray(grid, np.float32).reshape(-1, 2)
    return g[g[:, 0] >= 0]
