"""Shared input builders for the oracle (CPU) and GPU parity tests."""
import numpy as np

# Sampler grid of the Verify hand cases: r = 0.25 in the unit square, so the
# cell is 0.25 / sqrt(2) (FP32) and the grid is 6 x 6; r^2 = 0.0625 exactly.
VERIFY_R = 0.25
VERIFY_GW = 6


def verify_grid(samples):
    """A 6 x 6 sampler grid (x < 0: empty) holding `samples`, each given as ((cx, cy), (x, y))."""
    g = np.full((VERIFY_GW * VERIFY_GW, 2), -1.0, np.float32)
    for (cx, cy), p in samples:
        g[cy * VERIFY_GW + cx] = p
    return g


def verify_cases():
    """(name, grid, region, expected Verify result) for the reference's Verify (poisson_disk.cu:179-199):
    strict d^2 < r^2 in FP32, over the region expanded by two cells, each sample against its 5 x 5 cell
    neighbourhood."""
    f = np.float32
    just_above = float(np.nextafter(f(0.25), f(1)))
    full = (0, 0, VERIFY_GW, VERIFY_GW)
    cases = []
    # dx = 0.5 - 0.25000003 = 0.24999997 (exact), dx^2 < 0.0625: closer than r.
    g = verify_grid([((2, 1), (0.5, 0.3)), ((1, 1), (just_above, 0.3))])
    cases.append(("closer_than_r", g, full, False))
    # dx = 0.25 exactly: dx^2 == r^2 is not a conflict (strict test).
    g = verify_grid([((2, 1), (0.5, 0.3)), ((1, 1), (0.25, 0.3))])
    cases.append(("exactly_r", g, full, True))
    # A conflicting pair in cells (5, 4) and (5, 5): outside the window of region
    # (0, 0, 1, 1) expanded by two cells ([0, 3) x [0, 3)), so that check passes,
    # while the whole grid fails.
    g = verify_grid([((5, 4), (0.9, 0.8)), ((5, 5), (0.9, 0.95)), ((0, 0), (0.1, 0.1))])
    cases.append(("pair_outside_window", g, (0, 0, 1, 1), True))
    cases.append(("pair_outside_window_all", g, full, False))
    # One sample inside the window (cell (2, 2)), its conflicting partner just
    # outside it (cell (3, 2)): found through the inside sample's neighbourhood.
    g = verify_grid([((2, 2), (0.4, 0.4)), ((3, 2), (0.6, 0.4))])
    cases.append(("partner_outside_window", g, (0, 0, 1, 1), False))
    # The same two samples r + 0.01 apart: valid.
    g = verify_grid([((2, 2), (0.4, 0.4)), ((3, 2), (0.66, 0.4))])
    cases.append(("partner_far_enough", g, (0, 0, 1, 1), True))
    return cases


def plant_conflict(grid, gw, radius):
    """A copy of a valid sampler grid with one sample moved to 0.6 r from a neighbour in the next cell
    (returned with the index of the moved cell)."""
    g = grid.copy()
    live = np.nonzero(g[:, 0] >= 0)[0]
    for i in live:
        j = i + 1
        if j % gw == 0 or j >= len(g) or g[j, 0] < 0:
            continue
        g[j] = (g[i, 0] + np.float32(0.6 * radius), g[i, 1])
        return g, j
    raise RuntimeError("no neighbouring pair")


def clustered_sites(base, seed=5, centres=((0.37, 0.56), (0.64, 0.21), (0.5, 0.83)), per=40, spread=0.004):
    """`base` sites (a coarse Poisson-disk set) plus tight clusters of `per` sites within `spread` of each
    centre.  The cells of the coarse sites next to a cluster reach far beyond their 32 nearest sites (all in
    the cluster), so the clip cannot certify them from its 32-NN list and clips them against every site in
    range (the brute-force fallback)."""
    rng = np.random.default_rng(seed)
    parts = [np.asarray(base, np.float32)]
    for c in centres:
        a = rng.uniform(0, 2 * np.pi, per)
        rr = spread * np.sqrt(rng.uniform(0, 1, per))
        parts.append(np.stack([c[0] + rr * np.cos(a), c[1] + rr * np.sin(a)], 1).astype(np.float32))
    return np.concatenate(parts)
