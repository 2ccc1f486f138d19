"""Pins the CPU oracle (oracle/voro_oracle.cpp) before it is trusted as the checker.

Every function is checked against something other than its own formula:
  * pcg_hash / rand01            -- known-answer values (integer closed form)
  * k-NN ring search             -- brute force over all sites, several grid sizes, ties
  * VoroLoss + gradient          -- hand-computed case, FP64 central finite differences
  * Adam + MultiStepLR + clamp   -- textbook closed form of the first steps
  * cell clipping                -- scipy.spatial.Voronoi (Qhull) neighbours, area sum,
                                    neighbour symmetry, vertex equidistance
  * boundary assembly            -- two-site hand case, watertight star
  * Poisson dart thrower         -- determinism, region resample, yield; Verify on hand
                                    grids (conflicts closer than r, the r^2 boundary, the
                                    region window and its edge) and planted conflicts
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from cases import VERIFY_R, clustered_sites, plant_conflict, verify_cases

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _kat():
    with open(os.path.join(GOLDEN, "kat.json")) as f:
        return json.load(f)


def test_pcg_hash_kat():
    k = _kat()
    assert [O.pcg_hash(i) for i in range(4)] == [int(v, 16) for v in k["pcg_hash_0_3"]]


def test_rand01_kat():
    k = _kat()
    got = [O.rand01_loss(i, 1, 1) for i in range(4)]
    assert got == pytest.approx(k["rand01_loss_i0_3_iter1_seed1"], abs=5e-9)
    assert O.rand01_poisson(0, 0, 1, 0) == pytest.approx(k["rand01_poisson_c0_r0_s1"][0], abs=5e-9)
    assert O.rand01_poisson(0, 0, 1, 1) == pytest.approx(k["rand01_poisson_c0_r0_s1"][1], abs=5e-9)
    # 24-bit resolution: every value is a multiple of 2^-24 in [0, 1)
    vals = np.array([O.rand01_loss(i, 7, 3) for i in range(2000)])
    assert np.all(vals >= 0) and np.all(vals < 1)
    assert np.all(np.round(vals * 2 ** 24) == vals * 2 ** 24)


def test_pcg_hash_bit_formula():
    # An independent restatement in Python integers (u32 wrap-around).
    def h(v):
        state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
        word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
        return ((word >> 22) ^ word) & 0xFFFFFFFF
    for v in [0, 1, 17, 2 ** 31 - 1, 2 ** 32 - 1, 123456789]:
        assert O.pcg_hash(v) == h(v)


def test_loss_hand_case():
    """c0=(0,0), cj=(1,0), p=(0.25,0.3): d=-0.25, L=0.0625, dL/dcj=(0.25,-0.15), dL/dc0=(0.25,0.15)."""
    s = np.array([[0, 0], [1, 0]], np.float64)
    t = np.array([[0.25, 0.3]])
    r = O.loss_grad(s, t, 2, 2, 2, 1.0, real=1, per_target=True)
    assert r["loss_sum"] == pytest.approx(0.0625)
    assert r["c0"][0] == 0 and r["bestj"][0] == 1
    np.testing.assert_allclose(r["grad"], [[0.25, 0.15], [0.25, -0.15]], atol=1e-15)
    k = _kat()["loss_hand"]
    np.testing.assert_allclose(r["grad"], k["grad"], atol=1e-15)


def _loss64(sites, targets, K):
    return O.loss_grad(sites, targets, K, 1.0, 1.0, 0.25, real=1, per_target=True)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_gradient_finite_differences(seed):
    rng = np.random.default_rng(seed)
    sites = rng.uniform(0.1, 0.9, (6, 2))
    targets = rng.uniform(0.2, 0.8, (40, 2))
    K = 5
    base = _loss64(sites, targets, K)
    h = 1e-7
    fd = np.zeros_like(sites)
    for i in range(len(sites)):
        for c in range(2):
            sp, sm = sites.copy(), sites.copy()
            sp[i, c] += h
            sm[i, c] -= h
            rp, rm = _loss64(sp, targets, K), _loss64(sm, targets, K)
            # the discrete choices (k-NN set, c0, partner) must not change under +-h
            assert np.array_equal(rp["bestj"], base["bestj"]) and np.array_equal(rm["bestj"], base["bestj"])
            fd[i, c] = (rp["loss_sum"] - rm["loss_sum"]) / (2 * h)
    np.testing.assert_allclose(base["grad"], fd, rtol=1e-5, atol=1e-7)


def test_gradient_fp32_matches_fp64_closely():
    rng = np.random.default_rng(5)
    sites = rng.uniform(0, 1, (500, 2))
    targets = rng.uniform(0.1, 0.9, (4000, 2))
    g64 = O.loss_grad(sites, targets, 16, 1, 1, 0.1, real=1)["grad"]
    g32 = O.loss_grad(sites.astype(np.float32), targets.astype(np.float32), 16, 1, 1, 0.1, real=0)["grad"]
    assert np.linalg.norm(g32 - g64) / np.linalg.norm(g64) < 1e-5


@pytest.mark.parametrize("real", [0, 1])
@pytest.mark.parametrize("cell", [0.02, 0.05, 0.11, 0.5, 2.0])
def test_knn_ring_equals_brute_force(real, cell):
    rng = np.random.default_rng(11)
    sites = rng.uniform(0, 1, (700, 2))
    q = np.concatenate([rng.uniform(-0.1, 1.1, (600, 2)), sites[:50]])  # some outside the domain, some on sites
    for K in (1, 8, 16, 32):
        a = O.knn(sites, q, K, 1, 1, cell, real=real)
        b = O.knn(sites, q, K, 1, 1, cell, real=real, brute=True)
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1])


def test_knn_order_and_ties():
    """Exact ties (lattice) are ordered by index; lists are sorted by (d2, index)."""
    g = (np.arange(8) + 0.5) / 8
    sites = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2).astype(np.float32)
    q = np.array([[0.5, 0.5], [0.3125, 0.3125]], np.float32)  # equidistant to 4 sites
    idx, d2 = O.knn(sites, q, 12, 1, 1, 0.125)
    for row_i, row_d in zip(idx, d2):
        keys = list(zip(row_d.tolist(), row_i.tolist()))
        assert keys == sorted(keys)
    assert d2[0][0] == d2[0][1] == d2[0][2] == d2[0][3]
    assert list(idx[0][:4]) == sorted(idx[0][:4])
    # brute force agrees
    assert np.array_equal(idx, O.knn(sites, q, 12, 1, 1, 0.3, brute=True)[0])


def test_knn_fewer_sites_than_k():
    sites = np.array([[0.2, 0.2], [0.8, 0.8]], np.float32)
    idx, d2 = O.knn(sites, np.array([[0.5, 0.4]], np.float32), 4, 1, 1, 0.1)
    assert list(idx[0]) == [0, 1, -1, -1]
    assert d2[0][2] == np.finfo(np.float32).max


def test_subsampling_count_matches_hash():
    rng = np.random.default_rng(2)
    sites = rng.uniform(0, 1, (200, 2)).astype(np.float32)
    targets = rng.uniform(0, 1, (3000, 2)).astype(np.float32)
    r = O.loss_grad(sites, targets, 8, 1, 1, 0.1, it=3, fraction=0.2, seed=9)
    expect = sum(O.rand01_loss(i, 3, 9) < 0.2 for i in range(3000))
    assert r["count"] == expect


def test_empty_targets_are_skipped():
    sites = np.array([[0.2, 0.2], [0.8, 0.8], [0.5, 0.1]], np.float32)
    targets = np.array([[0.5, 0.5], [-1, -1], [np.nan, np.nan], [0.4, 0.6]], np.float32)
    r = O.loss_grad(sites, targets, 3, 1, 1, 0.1, per_target=True)
    assert r["count"] == 2
    assert list(r["c0"][[1, 2]]) == [-1, -1]


def test_adam_first_steps_closed_form():
    sc = O.make_sched(0.01, 1.0, 1.0, milestones=(1, 2, 100, 200), gamma=0.5)
    sites = np.array([[0.5, 0.5], [0.001, 0.999]], np.float32)
    m = np.zeros_like(sites)
    v = np.zeros_like(sites)
    grad = np.array([[3.0, -2.0], [1.0, -1.0]])  # summed over count=2 targets
    O.adam(sites, m, v, grad, 1, 2, sc)
    # step 1: m_hat/sqrt(v_hat) = g/|g| -> moves by lr*sign(g) (eps negligible); clamp at 0 and 1
    np.testing.assert_allclose(sites[0], [0.49, 0.51], atol=1e-7)
    np.testing.assert_allclose(sites[1], [0.0, 1.0], atol=0)
    np.testing.assert_allclose(m[0], [0.1 * 1.5, -0.1], rtol=1e-6)
    # (1 - 0.999) rounded to FP32 is 0.001 * (1 - 1.3e-5)
    np.testing.assert_allclose(v[0], [0.001 * 2.25, 0.001], rtol=2e-5)
    # step 2 (t-1 = 1 >= milestone 1): lr halved; same gradient -> moves by lr/2 again
    before = sites.copy()
    O.adam(sites, m, v, grad, 2, 2, sc)
    np.testing.assert_allclose(sites[0] - before[0], [-0.005, 0.005], atol=1e-7)
    # step 3 (t-1 = 2 >= milestones 1, 2): lr quartered
    before = sites.copy()
    O.adam(sites, m, v, grad, 3, 2, sc)
    np.testing.assert_allclose(sites[0] - before[0], [-0.0025, 0.0025], atol=1e-7)


def test_milestone_scaling_rounds_half_away():
    assert O.scale_milestones((80, 150, 200, 250), 100, 400) == [20, 38, 50, 63]
    assert O.scale_milestones((80, 150, 200, 250), 200, 400) == [40, 75, 100, 125]


def _clip_case(n=300, seed=4):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (n, 2)).astype(np.float32)


def test_clip_matches_qhull_neighbours_and_area():
    from scipy.spatial import Voronoi
    sites = _clip_case()
    c = O.clip(sites, 1.0, 1.0, 0.1)
    assert c["overflow"] == 0
    # area sum
    area = 0.0
    for i in range(len(sites)):
        v = c["verts"][i, :c["count"][i]].astype(np.float64)
        area += 0.5 * np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1])
    assert area == pytest.approx(1.0, abs=1e-5)
    # neighbour symmetry
    nb = [set(int(l) for l in c["labels"][i, :c["count"][i]] if l >= 0) for i in range(len(sites))]
    for i, s in enumerate(nb):
        for j in s:
            assert i in nb[j]
    # interior cells (not touching the box) have exactly Qhull's Voronoi neighbours
    vor = Voronoi(sites.astype(np.float64))
    qn = [set() for _ in sites]
    for a, b in vor.ridge_points:
        qn[a].add(int(b))
        qn[b].add(int(a))
    checked = 0
    for i in range(len(sites)):
        labels = c["labels"][i, :c["count"][i]]
        if (labels < 0).any():
            continue
        assert nb[i] == qn[i]
        checked += 1
    assert checked > 150


def test_clip_vertices_equidistant_to_label_triples():
    sites = _clip_case(200, 9)
    c = O.clip(sites, 1.0, 1.0, 0.1)
    s64 = sites.astype(np.float64)
    for i in range(len(sites)):
        n = c["count"][i]
        lab = c["labels"][i, :n]
        for k in range(n):
            a, b = lab[k - 1], lab[k]  # vertex k sits between edges k-1 and k
            v = c["verts"][i, k].astype(np.float64)
            di = np.hypot(*(v - s64[i]))
            for l in (a, b):
                if l >= 0:
                    assert abs(np.hypot(*(v - s64[l])) - di) < 2e-6


def test_clip_fp32_and_fp64_sites_same_topology():
    sites = _clip_case(400, 12)
    a = O.clip(sites, 1.0, 1.0, 0.1, real=0)
    b = O.clip(sites.astype(np.float64), 1.0, 1.0, 0.1, real=1)
    assert np.array_equal(a["count"], b["count"])
    assert np.array_equal(a["labels"], b["labels"])
    assert np.array_equal(a["verts"], b["verts"])


def test_clip_uncertified_cell_falls_back_to_brute_force():
    """With Kc=2 the k-NN list cannot certify most cells; the brute-force fallback still
    produces the same polygons as Kc=32."""
    sites = _clip_case(120, 3)
    a = O.clip(sites, 1.0, 1.0, 0.1, Kc=32)
    b = O.clip(sites, 1.0, 1.0, 0.1, Kc=2)
    assert b["unsafe"] > 0
    assert np.array_equal(a["count"], b["count"])
    for i in range(len(sites)):
        assert set(a["labels"][i, :a["count"][i]]) == set(b["labels"][i, :b["count"][i]])


def test_boundary_two_sites():
    sites = np.array([[0.25, 0.5], [0.75, 0.5]], np.float32)
    c = O.clip(sites, 1.0, 1.0, 0.5)
    inside = np.array([1, 0], np.uint8)
    b = O.boundary(c["verts"], c["labels"], c["count"], inside)
    assert b["watertight"] and b["loops"] == 1
    assert len(b["edges"]) == 4 and len(b["vertices"]) == 4
    xs = sorted(set(np.round(b["vertices"][:, 0], 6)))
    assert xs == [0.0, 0.5]


def test_boundary_star_watertight_one_loop():
    rng = np.random.default_rng(1)
    grid = O.poisson(1, 1, 0.03, 30, 1)
    sites = grid[grid[:, 0] >= 0]
    c = O.clip(sites, 1.0, 1.0, 0.1)
    # inside = site within a disc of radius 0.3: one loop
    inside = (np.hypot(sites[:, 0] - 0.5, sites[:, 1] - 0.5) < 0.3).astype(np.uint8)
    b = O.boundary(c["verts"], c["labels"], c["count"], inside)
    assert b["watertight"] and b["loops"] == 1
    # two separated discs: two loops
    inside2 = ((np.hypot(sites[:, 0] - 0.25, sites[:, 1] - 0.5) < 0.15) |
               (np.hypot(sites[:, 0] - 0.75, sites[:, 1] - 0.5) < 0.15)).astype(np.uint8)
    b2 = O.boundary(c["verts"], c["labels"], c["count"], inside2)
    assert b2["watertight"] and b2["loops"] == 2
    del rng


def test_poisson_valid_deterministic_and_region():
    g1 = O.poisson(1, 1, 0.02, 30, 5)
    g2 = O.poisson(1, 1, 0.02, 30, 5)
    assert np.array_equal(g1, g2)
    assert O.poisson_verify(1, 1, 0.02, g1)
    gw, gh, cs = O.poisson_dims(1, 1, 0.02)
    assert cs == np.float32(np.float32(0.02) / np.sqrt(np.float32(2)))
    g3 = O.poisson(1, 1, 0.02, 30, 5, region=(10, 10, 30, 25), round_base=30, grid=g1.copy())
    assert O.poisson_verify(1, 1, 0.02, g3)
    inside = np.zeros((gh, gw), bool)
    inside[10:25, 10:30] = True
    assert np.array_equal(g3[~inside.ravel()], g1[~inside.ravel()])
    assert not np.array_equal(g3[inside.ravel()], g1[inside.ravel()])
    # every sample lies in its own cell
    live = g1[:, 0] >= 0
    cells = np.nonzero(live)[0]
    assert np.all((g1[live, 0] / cs).astype(int) == cells % gw)


@pytest.mark.parametrize("name,grid,region,expected", verify_cases(), ids=[c[0] for c in verify_cases()])
def test_poisson_verify_hand_cases(name, grid, region, expected):
    assert O.poisson_verify(1, 1, VERIFY_R, grid, region=region) == expected


def test_poisson_verify_planted_conflict():
    """A valid sampled grid with one sample moved within r of its neighbour: invalid over the whole grid and
    over a region containing it; still valid over a region whose window (expanded by 2 cells) misses it."""
    r = 0.02
    g = O.poisson(1, 1, r, 30, 9)
    gw, gh, _ = O.poisson_dims(1, 1, r)
    assert O.poisson_verify(1, 1, r, g)
    bad, j = plant_conflict(g, gw, r)
    cx, cy = j % gw, j // gw
    assert not O.poisson_verify(1, 1, r, bad)
    assert not O.poisson_verify(1, 1, r, bad, region=(cx, cy, cx + 1, cy + 1))
    far = (0, 0, 1, 1) if cx > 8 or cy > 8 else (gw - 1, gh - 1, gw, gh)
    assert O.poisson_verify(1, 1, r, bad, region=far)


def test_clip_fallback_on_clustered_sites_matches_certified_clip():
    """Coarse Poisson sites beside tight clusters of 40 sites: the cells next to a cluster cannot be certified
    from their 32 nearest sites and take the brute-force fallback (unsafe > 0).  Their polygons have the same
    neighbours as a clip whose lists are long enough to certify every cell (Kc = 128, no fallback), vertices
    agree to 1e-6, and the areas sum to the domain's."""
    base = O.poisson(1, 1, 0.1, 30, 3)
    sites = clustered_sites(base[base[:, 0] >= 0])
    a = O.clip(sites, 1.0, 1.0, 0.2)
    b = O.clip(sites, 1.0, 1.0, 0.2, Kc=128)
    assert a["unsafe"] > 0 and b["unsafe"] == 0 and a["overflow"] == 0
    assert np.array_equal(a["count"], b["count"])
    area = 0.0
    for i in range(len(sites)):
        n = a["count"][i]
        assert set(a["labels"][i, :n]) == set(b["labels"][i, :n])
        assert np.abs(np.sort(a["verts"][i, :n], 0) - np.sort(b["verts"][i, :n], 0)).max() < 1e-6
        v = a["verts"][i, :n].astype(np.float64)
        area += 0.5 * np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1])
    assert area == pytest.approx(1.0, abs=1e-6)


def test_poisson_yield():
    ys = []
    for seed in range(3):
        g = O.poisson(1, 1, 0.01, 30, seed)
        ys.append((g[:, 0] >= 0).sum() * 0.01 ** 2)
    assert 0.66 < np.mean(ys) < 0.70


def test_clip_lattice_fp32_equals_fp64():
    """Jittered lattice (+-1e-7): FP32-stored sites clip exactly like the same values in FP64
    (the clip orders neighbours by FP64 distance), so the topology cannot depend on the mode."""
    rng = np.random.default_rng(8)
    g = (np.arange(24) + 0.5) / 24
    sites = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    sites = (sites + rng.uniform(-1e-7, 1e-7, sites.shape)).astype(np.float32)
    a = O.clip(sites, 1.0, 1.0, 1 / 12, real=0)
    b = O.clip(sites.astype(np.float64), 1.0, 1.0, 1 / 12, real=1)
    assert np.array_equal(a["count"], b["count"]) and np.array_equal(a["labels"], b["labels"])
    assert np.array_equal(a["verts"], b["verts"])
    assert a["overflow"] == 0


def test_oracle_reproduces_committed_config0_fixture():
    """tests/golden/config0_small.npz (Poisson sites r = 0.05 seed 0, 1,024 star targets, k = 16; written by
    make_golden.py) is reproduced bit for bit: sampler grid, brute-force k-NN lists, per-target terms, gradient."""
    z = np.load(os.path.join(GOLDEN, "config0_small.npz"))
    assert np.array_equal(O.poisson(1.0, 1.0, 0.05, 30, 0), z["grid"])
    sites, targets = z["sites"], z["targets"]
    idx, d2 = O.knn(sites, targets, 16, 1.0, 1.0, 0.1, brute=False)
    assert np.array_equal(idx, z["knn"]) and np.array_equal(d2, z["knn_d2"])
    r = O.loss_grad(sites, targets, 16, 1.0, 1.0, 0.1, per_target=True)
    for key in ("c0", "bestj", "loss_t", "grad"):
        assert np.array_equal(r[key], z[key]), key
