"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars: k-NN lists, nearest site / bisector partner, per-target loss terms,
clipped cells and Poisson grids bit-exact against the oracle's FP32 semantics;
FP32 gradient sums (64-bit fixed point, order-independent) and whole optimizer
trajectories bit-exact against the oracle; gradients within 1e-5 relative
(norm-wise) of the FP64 oracle; the FP64 mode within 1e-12 of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
from cases import VERIFY_R, clustered_sites, plant_conflict, verify_cases

vmhip = pytest.importorskip("vmhip")
from vmhip import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

W = H = 1.0


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


def _config0(r=0.025, n_targets=8192, seed=0):
    grid = O.poisson(W, H, r, 30, seed)
    gw, gh, cs = O.poisson_dims(W, H, r)
    return grid, (gw, gh, cs), wl.star(n_targets, seed=seed)


def _optimizer(r, K, precision=0, iterations=200, fraction=1.0, grid_cell=0.0, rings=0):
    return vmhip.Optimizer(W, H, r, loss_k=K, sample_fraction=fraction, lr_per_spacing=1e-3 / r,
                           iterations=iterations, precision=precision, grid_cell=grid_cell, stage_rings=rings)


def _lr(r):
    """lr exactly as the optimizer forms it: float(lr_per_spacing) * float(spacing) in FP32."""
    return float(np.float32(np.float32(1e-3 / r) * np.float32(r)))


def _box(r, gw, gh, cs):
    """Target box of a full-rectangle schedule, in FP32 as the optimizer forms it."""
    f = np.float32
    m = f(2) * f(r)
    return (float(f(0) * f(cs) - m), float(f(0) * f(cs) - m), float(f(gw) * f(cs) + m), float(f(gh) * f(cs) + m))


def _slots(grid):
    return np.nonzero(grid[:, 0] >= 0)[0]


def _map(idx, slot_of):
    """Instance-local site indices -> slot numbers; negative labels (missing, box sides) pass through."""
    idx = np.asarray(idx)
    return np.where(idx >= 0, slot_of[np.maximum(idx, 0)], idx)


@pytest.mark.parametrize("K", [12, 16, 32])
def test_knn_query_bitexact_config0(K):
    r = 0.025
    grid, (gw, gh, cs), targets = _config0(r)
    opt = _optimizer(r, K)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    sites = wl.live_sites(grid)
    idx, d2 = opt.knn(targets, K)
    ridx, rd2 = O.knn(sites, targets, K, W, H, 0.1, brute=True)
    assert np.array_equal(idx, _map(ridx, _slots(grid)))
    assert np.array_equal(d2, rd2)


@pytest.mark.parametrize("exact", [False, True], ids=["screened", "exact"])
@pytest.mark.parametrize("cell_factor,rings,K,r", [(2.0, 2, 16, 0.025), (1.0, 1, 16, 0.025), (3.0, 1, 16, 0.025),
                                                   (0.5, 2, 16, 0.025), (0.0, 0, 16, 0.025), (0.0, 0, 12, 0.025),
                                                   (0.0, 0, 8, 0.01), (0.0, 0, 24, 0.01), (0.0, 0, 16, 0.006),
                                                   (0.0, 0, 2, 0.025), (0.0, 0, 32, 0.01)])
def test_loss_kernel_terms_bitexact(cell_factor, rings, K, r, exact, monkeypatch):
    """k-NN inside the loss kernel, nearest site, partner and loss term, for several grid shapes
    (cell_factor 0: the tuner's choice, which exercises the wavefront bounding-box path), through
    the screened FP32 term and through the exact selection alone (VM_LOSS_EXACT=1)."""
    if exact:
        monkeypatch.setenv("VM_LOSS_EXACT", "1")
    grid, (gw, gh, cs), targets = _config0(r, n_targets=8192 if r > 0.01 else 65536)
    sites = wl.live_sites(grid)
    spacing = (1.0 / len(sites)) ** 0.5
    opt = _optimizer(r, K, grid_cell=cell_factor * spacing, rings=rings)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    d = opt.loss_pass(len(targets))
    ref = O.loss_grad(sites, targets, K, W, H, 0.1, per_target=True)
    ridx, _ = O.knn(sites, targets, K, W, H, 0.1)
    so = _slots(grid)
    assert np.array_equal(d["knn"], _map(ridx, so))
    assert np.array_equal(d["nearest"], _map(ref["c0"], so))
    assert np.array_equal(d["partner"], _map(ref["bestj"], so))
    assert np.array_equal(d["loss_terms"], ref["loss_t"].astype(np.float64))
    assert d["count"] == ref["count"]
    assert d["loss_sum"] == pytest.approx(ref["loss_sum"], rel=1e-12)


def test_loss_terms_coincident_sites():
    """Sites placed exactly on top of others (zero-length bisector normals, skipped by the reference's
    l < 1e-12 test) and a target exactly on a site: terms, partners and gradients stay bit-exact."""
    r, K = 0.025, 16
    grid, (gw, gh, cs), targets = _config0(r)
    so = _slots(grid)
    g = grid.copy()
    for a, b in [(10, 11), (200, 201), (500, 520), (700, 650)]:
        g[so[b]] = g[so[a]]
    targets = targets.copy()
    targets[:64] = g[so[300:364]]
    sites = wl.live_sites(g)
    opt = _optimizer(r, K)
    opt.set_targets(targets)
    opt.set_generators(g, gw, gh, cs)
    d = opt.loss_pass(len(targets))
    ref = O.loss_grad(sites, targets, K, W, H, 0.1, per_target=True)
    assert np.array_equal(d["nearest"], _map(ref["c0"], so))
    assert np.array_equal(d["partner"], _map(ref["bestj"], so))
    assert np.array_equal(d["loss_terms"], ref["loss_t"].astype(np.float64))
    assert np.array_equal(d["grad"][so], ref["grad"])


def test_batch_screening_resolves_most_targets():
    """On random-polygon fits the screened FP32 term settles nearly every target; the exact selection
    runs for the rest (counted by the kernel), and the trajectories stay bit-exact (tested elsewhere)."""
    b = vmhip.Batch()
    n_t = 0
    for i in range(6):
        grid = O.poisson(W, H, 0.012 * (1 + 0.3 * i), 30, 40 + i)
        sites = wl.live_sites(grid)
        tg = wl.polygon(8 * len(sites), seed=60 + i)
        n_t += len(tg)
        b.add(W, H, sites, tg, 1e-3, loss_k=16)
    b.commit()
    st = b.loss_stats()
    print(st)
    assert st["fast_waves"] > 0.95 * (st["fast_waves"] + st["ring_staged_waves"] + st["ring_global_waves"])
    assert st["exact_fallback_targets"] < 0.01 * n_t


def test_knn_lattice_ties_bitexact():
    """Jittered lattice (+-1e-7): many exactly tied squared distances; order is (d2, index)."""
    sites = wl.lattice(64, 1e-7, seed=3)
    targets = wl.circle(16384, radius=0.3, sigma=1e-4, seed=3)
    # Sites as slots of a 64x64 slot grid, all live.
    opt = _optimizer(1 / 64, 16)
    opt.set_targets(targets)
    opt.set_generators(sites, 64, 64, 1 / 64)
    idx, d2 = opt.knn(targets, 16)
    ridx, rd2 = O.knn(sites, targets, 16, W, H, 0.1, brute=True)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(d2, rd2)
    # Query the sites themselves: exact ties between lattice neighbours.
    idx2, _ = opt.knn(sites, 9)
    ridx2, _ = O.knn(sites, sites, 9, W, H, 0.1, brute=True)
    assert np.array_equal(idx2, ridx2)
    d = opt.loss_pass(len(targets))
    ref = O.loss_grad(sites, targets, 16, W, H, 0.1, per_target=True)
    assert np.array_equal(d["partner"], ref["bestj"])
    assert np.array_equal(d["loss_terms"], ref["loss_t"].astype(np.float64))


def test_gradient_vs_fp64_oracle():
    r, K = 0.025, 16
    grid, (gw, gh, cs), targets = _config0(r)
    sites = wl.live_sites(grid)
    opt = _optimizer(r, K)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    g = opt.loss_pass(len(targets), want_terms=False)["grad"][_slots(grid)]
    r32 = O.loss_grad(sites, targets, K, W, H, 0.1)
    r64 = O.loss_grad(sites.astype(np.float64), targets.astype(np.float64), K, W, H, 0.1, real=1)
    e32 = np.linalg.norm(g - r32["grad"]) / np.linalg.norm(r32["grad"])
    e64 = np.linalg.norm(g - r64["grad"]) / np.linalg.norm(r64["grad"])
    print(f"grad rel err vs FP32 oracle {e32:.3e}, vs FP64 oracle {e64:.3e} (max abs {np.abs(g - r64['grad']).max():.3e})")
    assert np.array_equal(g, r32["grad"])
    assert e64 < 1e-5


def test_one_step_matches_oracle():
    r, K = 0.025, 16
    grid, (gw, gh, cs), targets = _config0(r)
    sites = wl.live_sites(grid)
    opt = _optimizer(r, K)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    opt.step(1)
    got = opt.generators()[_slots(grid)]
    sc = O.make_sched(_lr(r), W, H)
    ref, _, _, lh, ch = O.fit(sites, targets, K, 0.1, sc, 1, box=_box(r, gw, gh, cs))
    assert np.array_equal(got, ref)
    assert opt.state()["iteration"] == 1


def test_subsampled_steps_count_exact():
    """With sample_fraction 0.2 the per-step target count is an exact integer parity check."""
    r, K = 0.025, 12
    grid, (gw, gh, cs), targets = _config0(r)
    sites = wl.live_sites(grid)
    opt = _optimizer(r, K, fraction=0.2, iterations=400)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    opt.step(1)
    sc = O.make_sched(_lr(r), W, H)
    ref, _, _, lh, ch = O.fit(sites, targets, K, 0.1, sc, 3, fraction=0.2, seed=1, box=_box(r, gw, gh, cs))
    opt.step(2)
    st = opt.state()
    assert np.array_equal(opt.generators()[_slots(grid)], ref)
    assert st["last_batch_loss"] == pytest.approx(lh[-1] / ch[-1], rel=1e-6)


def test_fit_config0_final_loss_close_to_oracle():
    r, K, iters = 0.025, 16, 200
    grid, (gw, gh, cs), targets = _config0(r)
    sites = wl.live_sites(grid)
    opt = _optimizer(r, K, iterations=iters)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    opt.optimize()
    gpu_loss = opt.evaluate_loss()
    sc = O.make_sched(_lr(r), W, H)
    ref, _, _, lh, ch = O.fit(sites, targets, K, 0.1, sc, iters, box=_box(r, gw, gh, cs))
    # The whole 200-step trajectory is bit-exact (exact k-NN, identical per-target
    # terms, order-independent gradient sums, identical Adam arithmetic).
    assert np.array_equal(opt.generators()[_slots(grid)], ref)
    r_loss = O.loss_grad(ref, targets, K, W, H, 0.1)
    ref_mean = r_loss["loss_sum"] / r_loss["count"]
    l0 = O.loss_grad(sites, targets, K, W, H, 0.1)
    print(f"initial {l0['loss_sum'] / l0['count']:.4e}  gpu {gpu_loss:.4e}  oracle {ref_mean:.4e}")
    assert gpu_loss < 0.5 * l0["loss_sum"] / l0["count"]
    assert gpu_loss == pytest.approx(ref_mean, rel=1e-12)


def test_clip_bitexact_and_watertight():
    r, K = 0.025, 16
    grid, (gw, gh, cs), targets = _config0(r)
    sites = wl.live_sites(grid)
    opt = _optimizer(r, K)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    opt.step(20)
    moved = opt.generators()
    so = _slots(grid)
    live = moved[so]
    # occupancy: voxel centre inside the star polygon (64x64 voxels of 1/64)
    masks, tx, ty, vox = _star_masks(64)
    opt.set_occupancy(masks, tx, ty, vox)
    mesh = opt.extract()
    ref = O.clip(live, W, H, 0.1, masks=masks, tiles=(tx, ty), voxel=vox)
    cnt = mesh["cell_start"][1:] - mesh["cell_start"][:-1]
    assert np.array_equal(cnt[so], ref["count"])
    for k, s in enumerate(so):
        b = mesh["cell_start"][s]
        n = ref["count"][k]
        assert np.array_equal(mesh["cell_verts"][b:b + n], ref["verts"][k, :n])
        assert np.array_equal(mesh["cell_labels"][b:b + n], _map(ref["labels"][k, :n], so) if n else [])
    assert np.array_equal(mesh["inside"][so], ref["inside"])
    assert mesh["unsafe_cells"] == ref["unsafe"] and mesh["overflow_cells"] == ref["overflow"] == 0
    b = O.boundary(ref["verts"], ref["labels"], ref["count"], ref["inside"])
    assert mesh["watertight"] and b["watertight"]
    assert mesh["loops"] == b["loops"] >= 1
    assert len(mesh["vertices"]) == len(b["vertices"]) and len(mesh["edges"]) == len(b["edges"])
    area = 0.0
    for k in range(len(so)):
        v = ref["verts"][k, :ref["count"][k]].astype(np.float64)
        area += 0.5 * np.sum(v[:, 0] * np.roll(v[:, 1], -1) - np.roll(v[:, 0], -1) * v[:, 1])
    assert area == pytest.approx(1.0, abs=1e-5)


def _star_masks(vox_n):
    """Occupancy of the config-0 star: bit set when the voxel centre is inside (even-odd)."""
    import math
    ang = math.pi / 2 + np.arange(10) * math.pi / 5
    rad = np.where(np.arange(10) % 2 == 0, 0.40, 0.18)
    poly = np.stack([0.5 + rad * np.cos(ang), 0.5 + rad * np.sin(ang)], axis=1)
    vox = 1.0 / vox_n
    tx = ty = (vox_n + 7) // 8
    masks = np.zeros(tx * ty, np.uint64)
    for j in range(vox_n):
        for i in range(vox_n):
            x, y = (i + 0.5) * vox, (j + 0.5) * vox
            inside = False
            for s in range(10):
                a, b = poly[s], poly[s - 1]
                if (a[1] > y) != (b[1] > y) and x < a[0] + (y - a[1]) * (b[0] - a[0]) / (b[1] - a[1]):
                    inside = not inside
            if inside:
                masks[(j >> 3) * tx + (i >> 3)] |= np.uint64(1) << np.uint64(((j & 7) << 3) | (i & 7))
    return masks, tx, ty, np.float32(vox)


@pytest.mark.parametrize("r,seed", [(0.025, 0), (0.01, 7)])
def test_poisson_sampler_bitexact(r, seed):
    s = vmhip.Sampler(W, H, r, 30, seed)
    ref = O.poisson(W, H, r, 30, seed)
    assert np.array_equal(s.cells(), ref)
    assert s.verify()
    # Region resample continues the round counter (fresh darts).
    s.resample(3, 4, 20, 17)
    ref = O.poisson(W, H, r, 30, seed, region=(3, 4, 20, 17), round_base=30, grid=ref)
    assert np.array_equal(s.cells(), ref)
    assert s.verify()


def test_sampler_verify_reports_conflicts():
    """The product's Verify (vm_sampler_verify) on planted grids, against the oracle and the hand-derived
    expectations: conflicts closer than r, the strict r^2 boundary, the region window and its edge, and a
    conflict planted into a sampled grid (invalid over it, valid over a window that misses it)."""
    s = vmhip.Sampler(W, H, VERIFY_R, 30, 1)
    assert (s.grid_w, s.grid_h) == (6, 6)
    for name, grid, region, expected in verify_cases():
        s.load_cells(grid)
        assert np.array_equal(s.cells(), grid)
        got = s.verify(*region)
        assert got == expected == O.poisson_verify(W, H, VERIFY_R, grid, region=region), name
    r = 0.02
    s2 = vmhip.Sampler(W, H, r, 30, 9)
    g = s2.cells()
    assert s2.verify() and np.array_equal(g, O.poisson(W, H, r, 30, 9))
    bad, j = plant_conflict(g, s2.grid_w, r)
    s2.load_cells(bad)
    cx, cy = j % s2.grid_w, j // s2.grid_w
    assert not s2.verify() and not s2.verify(cx, cy, cx + 1, cy + 1)
    far = (0, 0, 1, 1) if cx > 8 or cy > 8 else (s2.grid_w - 1, s2.grid_h - 1, s2.grid_w, s2.grid_h)
    assert s2.verify(*far)
    # A resample after the load continues from the loaded grid: the region is redrawn (valid again there),
    # the rest of the grid is kept.
    reg = (max(cx - 1, 0), max(cy - 1, 0), min(cx + 2, s2.grid_w), min(cy + 2, s2.grid_h))
    s2.resample(*reg)
    ref = O.poisson(W, H, r, 30, 9, region=reg, round_base=30, grid=bad.copy())
    assert np.array_equal(s2.cells(), ref)
    assert s2.verify()


def test_clip_fallback_runs_on_gpu_bitexact():
    """Coarse Poisson sites beside tight clusters: the cells next to a cluster cannot be certified from their 32
    nearest sites, so the clip kernel's brute-force fallback runs (unsafe cells > 0, the same count as the
    oracle's); every cell is bit-exact against the oracle's clip, in both precision modes."""
    base = O.poisson(W, H, 0.1, 30, 3)
    sites = clustered_sites(base[base[:, 0] >= 0])
    ref = O.clip(sites, W, H, 0.2)
    assert ref["unsafe"] > 0
    for precision in (0, 1):
        print(f"case: clip fallback, precision {precision}", flush=True)
        b = vmhip.Batch(precision)
        b.add(W, H, sites, wl.star(8 * len(sites), seed=3), 1e-3, loss_k=16)
        b.commit()
        m = b.extract(0)
        assert m["unsafe_cells"] == ref["unsafe"] and m["overflow_cells"] == 0
        cnt = m["cell_start"][1:] - m["cell_start"][:-1]
        assert np.array_equal(cnt, ref["count"])
        for k in range(len(sites)):
            s0, n = m["cell_start"][k], ref["count"][k]
            assert np.array_equal(m["cell_verts"][s0:s0 + n], ref["verts"][k, :n])
            assert np.array_equal(m["cell_labels"][s0:s0 + n], ref["labels"][k, :n])
        b.close()


def test_poisson_batch_matches_single():
    radii = [0.05, 0.02, 0.03]
    grids, dims, _ = vmhip.poisson_batch([1, 1, 1], [1, 1, 1], radii, [5, 6, 7])
    for g, r, sd in zip(grids, radii, [5, 6, 7]):
        assert np.array_equal(g, O.poisson(W, H, r, 30, sd))


def test_fp64_mode_vs_fp64_oracle():
    r, K = 0.025, 16
    grid, (gw, gh, cs), targets = _config0(r)
    sites = wl.live_sites(grid)
    opt = _optimizer(r, K, precision=1)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    d = opt.loss_pass(len(targets))
    r64 = O.loss_grad(sites.astype(np.float64), targets.astype(np.float64), K, W, H, 0.1, real=1,
                      per_target=True)
    so = _slots(grid)
    assert np.array_equal(d["partner"], _map(r64["bestj"], so))
    assert np.array_equal(d["loss_terms"], r64["loss_t"])
    g = d["grad"][so]
    assert np.linalg.norm(g - r64["grad"]) / np.linalg.norm(r64["grad"]) < 1e-12


def test_batch_instances_match_plain():
    """A fit in a batch advances like the same fit alone (within FP32 atomic-order noise)."""
    r, K, iters = 0.05, 16, 30
    b = vmhip.Batch()
    cases = []
    for i in range(4):
        grid = O.poisson(W, H, r * (1 + 0.2 * i), 30, 10 + i)
        sites = wl.live_sites(grid)
        tg = wl.polygon(8 * len(sites), seed=20 + i)
        cases.append((sites, tg))
        b.add(W, H, sites, tg, 1e-3, loss_k=K)
    b.commit()
    b.step(iters)
    for i, (sites, tg) in enumerate(cases):
        sc = O.make_sched(1e-3, W, H)
        ref, _, _, lh, ch = O.fit(sites, tg, K, 0.1, sc, iters)
        assert np.array_equal(b.sites(i), ref)
        st = b.status(i)
        assert st["iteration"] == iters
        assert st["last_mean_loss"] == pytest.approx(lh[-1] / ch[-1], rel=1e-12)


@pytest.mark.parametrize("K,tpl", [(8, 2), (8, 4), (12, 2), (12, 4), (16, 2), (16, 4)])
def test_batch_targets_per_lane_bitexact(K, tpl):
    """The loss kernel's other chunk shapes (2 and 4 targets per lane, chunks of up to 128 and 256 targets):
    15 batched iterations of three fits bit-exact against the oracle."""
    r, iters = 0.04, 15
    b = vmhip.Batch()
    cases = []
    for i in range(3):
        grid = O.poisson(W, H, r * (1 + 0.25 * i), 30, 70 + i)
        sites = wl.live_sites(grid)
        tg = wl.polygon(8 * len(sites), seed=80 + i)
        cases.append((sites, tg))
        b.add(W, H, sites, tg, 1e-3, loss_k=K, targets_per_lane=tpl)
    b.commit()
    assert all(b.choice(i)["targets_per_lane"] == tpl for i in range(3))
    b.step(iters)
    for i, (sites, tg) in enumerate(cases):
        ref, _, _, _, _ = O.fit(sites, tg, K, 0.1, O.make_sched(1e-3, W, H), iters)
        assert np.array_equal(b.sites(i), ref), f"fit {i}"


@pytest.mark.parametrize("precision", [0, 1])
def test_clip_lattice_topology_zero_mismatches(precision):
    """Precision stress: jittered 128x128 lattice; clipped cells (labels and vertices) equal the FP64
    oracle's in both the FP32 and the all-FP64 mode."""
    sites = wl.lattice(128, 1e-7, seed=3)
    opt = _optimizer(1 / 128, 16, precision=precision)
    opt.set_targets(wl.circle(4096, radius=0.3, sigma=1e-4, seed=3))
    opt.set_generators(sites, 128, 128, 1 / 128)
    mesh = opt.extract()
    ref = O.clip(sites.astype(np.float64), W, H, 1 / 64, real=1)
    cnt = mesh["cell_start"][1:] - mesh["cell_start"][:-1]
    assert np.array_equal(cnt, ref["count"])
    mism = 0
    for i in range(len(sites)):
        b = mesh["cell_start"][i]
        n = cnt[i]
        mism += not (np.array_equal(mesh["cell_labels"][b:b + n], ref["labels"][i, :n]) and
                     np.array_equal(mesh["cell_verts"][b:b + n], ref["verts"][i, :n]))
    assert mism == 0
    assert mesh["overflow_cells"] == 0


def test_batch_extract_matches_oracle_clip_and_boundary():
    """vm_batch_extract: every fit of a batch clipped (FP64) and its boundary assembled, bit-exact against the
    oracle's clip and boundary assembly in the fit's site order; without occupancy no cell is inside."""
    r, K, iters = 0.04, 16, 15
    masks, tx, ty, vox = _star_masks(64)
    b = vmhip.Batch()
    for i in range(3):
        grid = O.poisson(W, H, r * (1 + 0.25 * i), 30, 30 + i)
        sites = wl.live_sites(grid)
        b.add(W, H, sites, wl.star(8 * len(sites), seed=40 + i), 1e-3, loss_k=K)
    b.commit()
    b.step(iters)
    for i in range(3):
        print(f"case: batch extract fit {i}", flush=True)
        live = b.sites(i)
        m = b.extract(i, masks, (tx, ty), vox)
        ref = O.clip(live, W, H, 0.1, masks=masks, tiles=(tx, ty), voxel=vox)
        assert np.array_equal(m["generators"], live)
        cnt = m["cell_start"][1:] - m["cell_start"][:-1]
        assert np.array_equal(cnt, ref["count"])
        for k in range(len(live)):
            s0, n = m["cell_start"][k], ref["count"][k]
            assert np.array_equal(m["cell_verts"][s0:s0 + n], ref["verts"][k, :n])
            assert np.array_equal(m["cell_labels"][s0:s0 + n], ref["labels"][k, :n])
        assert np.array_equal(m["inside"], ref["inside"])
        assert m["unsafe_cells"] == ref["unsafe"] and m["overflow_cells"] == ref["overflow"] == 0
        bd = O.boundary(ref["verts"], ref["labels"], ref["count"], ref["inside"])
        assert np.array_equal(m["vertices"], bd["vertices"]) and np.array_equal(m["edges"], bd["edges"])
        assert m["watertight"] and bd["watertight"] and m["loops"] == bd["loops"] >= 1
        bare = b.extract(i)
        assert not bare["inside"].any() and len(bare["edges"]) == 0
        assert np.array_equal(bare["cell_verts"], m["cell_verts"])


def _star_polygon(tips=5, center=(0.5, 0.5), r_out=0.40, r_in=0.18):
    k = np.arange(2 * tips)
    a = np.pi / 2 + k * np.pi / tips
    rad = np.where(k % 2 == 0, r_out, r_in)
    return np.stack([center[0] + rad * np.cos(a), center[1] + rad * np.sin(a)], 1).astype(np.float32)


def _live_targets(slots):
    """Canvas slot array -> (live targets in slot order, their slot numbers), as the optimizer compacts them."""
    live = np.nonzero(slots[:, 0] >= 0)[0]
    return slots[live], live.astype(np.int32)


def test_local_edit_matches_oracle():
    """The local edit sequence -- Canvas.StampDisc -> UpdateTargets(dirty) -> Sampler.Resample(rect) ->
    UpdateGenerators(rect) -> BeginSchedule(n, rect) -> n steps -- bit-exact against the oracle replaying it:
    the canvas's live slots as targets (subsampled by slot number), the resampled rect's generators in place of
    the fitted ones, Adam moments zeroed and sites movable inside the rect only, milestones scaled to the local
    schedule, the rect's target box and step numbers restarting at 1.  The first edit grows the canvas past the
    optimizer's target capacity (full re-upload); the second stays within it (dirty slots only)."""
    r, K, iters, local = 0.025, 12, 120, 50
    f = np.float32
    canvas = vmhip.Canvas(W, H, 1.0 / 1024)
    canvas.rasterize(_star_polygon())
    sampler = vmhip.Sampler(W, H, r, 30, 1)
    gw, gh, cs = sampler.grid_w, sampler.grid_h, sampler.cell_size
    grid = sampler.cells()
    slots = canvas.target_slots()
    opt = _optimizer(r, K, fraction=0.2, iterations=iters)
    opt.set_targets(slots)
    opt.set_generators(grid, gw, gh, cs)
    print("case: local edit, global fit", flush=True)
    opt.step(iters)
    tgt, tslot = _live_targets(slots)
    live = _slots(grid)
    ref, m, v, _, _ = O.fit(grid[live], tgt, K, 0.1, O.make_sched(_lr(r), W, H), iters, fraction=0.2, seed=1,
                            box=_box(r, gw, gh, cs), slot_of=tslot)
    assert np.array_equal(opt.generators()[live], ref)
    pos, mom, vel = grid.copy(), np.zeros_like(grid), np.zeros_like(grid)
    pos[live], mom[live], vel[live] = ref, m, v
    x, y = np.arange(gw * gh) % gw, np.arange(gw * gh) // gw
    sched = O.make_sched(_lr(r), W, H, milestones=O.scale_milestones((80, 150, 200, 250), local, iters))
    capacity = len(slots)
    for e, (c, brush, add) in enumerate((((0.5, 0.9), 0.05, True), ((0.5, 0.32), 0.04, False))):
        edit, dirty = canvas.stamp_disc(c, brush, add)
        slots = canvas.target_slots()
        assert edit["removed"] > 0 and edit["added"] > 0 and len(dirty) > 0
        assert (len(slots) > capacity) == (e == 0)
        capacity = max(capacity, len(slots) * 3 // 2, len(slots) + 64)
        opt.update_targets(dirty, slots)
        margin = brush + 2 * r
        rect = (max(int(np.floor((c[0] - margin) / cs)), 0), max(int(np.floor((c[1] - margin) / cs)), 0),
                min(int(np.ceil((c[0] + margin) / cs)), gw), min(int(np.ceil((c[1] + margin) / cs)), gh))
        print(f"case: local edit {e}, rect {rect}, {len(dirty)} dirty slots", flush=True)
        sampler.resample(*rect)
        cells = sampler.cells()
        opt.update_generators(cells, *rect)
        opt.begin_schedule(local, rect)
        opt.step(local)
        in_rect = (x >= rect[0]) & (x < rect[2]) & (y >= rect[1]) & (y < rect[3])
        pos[in_rect], mom[in_rect], vel[in_rect] = cells[in_rect], 0, 0
        live = np.nonzero(pos[:, 0] >= 0)[0]
        tgt, tslot = _live_targets(slots)
        box = tuple(float(f(a) * f(cs) + s * f(2) * f(r)) for a, s in zip(rect, (-1, -1, 1, 1)))
        ref, m, v, _, ch = O.fit(pos[live], tgt, K, 0.1, sched, local, fraction=0.2, seed=1, box=box,
                                 movable=in_rect[live], slot_of=tslot, m=mom[live], v=vel[live])
        got = opt.generators()
        assert ch.min() > 0
        assert not np.array_equal(ref[in_rect[live]], pos[live][in_rect[live]])  # the rect's sites moved
        assert np.array_equal(got[live], ref)
        pos[live], mom[live], vel[live] = ref, m, v


def _disc_grid_masks(vox_n, n_side, radius):
    """Occupancy of an n_side x n_side grid of discs (voxel centres strictly inside a disc)."""
    vox = 1.0 / vox_n
    tx = ty = (vox_n + 7) // 8
    c = (np.arange(vox_n) + 0.5) * vox
    X, Y = np.meshgrid(c, c)
    cx, cy = (np.floor(X * n_side) + 0.5) / n_side, (np.floor(Y * n_side) + 0.5) / n_side
    j, i = np.nonzero((X - cx) ** 2 + (Y - cy) ** 2 < radius ** 2)
    masks = np.zeros(tx * ty, np.uint64)
    np.bitwise_or.at(masks, (j >> 3) * tx + (i >> 3), np.left_shift(np.uint64(1), (((j & 7) << 3) | (i & 7)).astype(np.uint64)))
    return masks, tx, ty, np.float32(vox)


def test_boundary_long_bitexact_and_regrow():
    """A boundary of 576 discs (about 9,000 edges): more end points than the assembly's first capacity (the
    buffers grow and the assembly runs again) and more vertices than its LDS pointer jumping holds (the
    rounds run in global memory).  Vertices, edges, watertightness and loops are bit-exact against the
    oracle's sequential assembly, and a second extract (grown buffers) returns the same mesh."""
    n_side = 24
    masks, tx, ty, vox = _disc_grid_masks(1024, n_side, 0.3 / n_side)
    grid = O.poisson(W, H, wl.poisson_radius_for(40000), 30, 11)
    sites = wl.live_sites(grid)
    b = vmhip.Batch()
    b.add(W, H, sites, wl.star(8 * len(sites), seed=11), 1e-3, loss_k=16)
    b.commit()
    b.step(3)
    live = b.sites(0)
    print(f"case: long boundary, {len(live)} sites", flush=True)
    m = b.extract(0, masks, (tx, ty), vox)
    ref = O.clip(live, W, H, 0.05, masks=masks, tiles=(tx, ty), voxel=vox)
    assert np.array_equal(m["inside"], ref["inside"])
    bd = O.boundary(ref["verts"], ref["labels"], ref["count"], ref["inside"])
    assert len(bd["edges"]) > 4096 and len(bd["vertices"]) > 4096
    assert np.array_equal(m["vertices"], bd["vertices"]) and np.array_equal(m["edges"], bd["edges"])
    assert m["watertight"] == bd["watertight"] and m["loops"] == bd["loops"]
    assert bd["watertight"] and bd["loops"] >= n_side * n_side // 2
    again = b.extract(0, masks, (tx, ty), vox)
    assert np.array_equal(again["vertices"], m["vertices"]) and np.array_equal(again["edges"], m["edges"])
    assert again["loops"] == m["loops"]


def _grad_from_terms(res, mask, S):
    """Per-site gradient sums (float64) from an oracle pass's per-target terms, over the targets in mask."""
    c0, bj, tg = res["c0"][mask], res["bestj"][mask], res["term_grad"][mask].astype(np.float64)
    gx = np.bincount(bj, tg[:, 2], S) + np.bincount(c0, tg[:, 0], S)
    gy = np.bincount(bj, tg[:, 3], S) + np.bincount(c0, tg[:, 1], S)
    return np.stack([gx, gy], 1)


def test_large_instance_full_size_bitexact():
    """The large single instance at full size (Poisson sites for 262,144, 4,194,304 targets on 64 ellipses, k=16).
    FP32: the gradient of a loss pass is bit-identical to the oracle's, and three optimizer steps give bit-identical
    positions.  Against FP64: the FP32 selections (nearest site, bisector partner) agree with the FP64 ones except at
    near-ties below FP32 resolution (one target here); over the agreeing targets the gradient is within 1e-5
    (relative, norm-wise) of FP64, while the flipped terms dominate the whole-gradient figure, which is printed.  The
    FP64 mode is within 1e-12 of the FP64 oracle."""
    K = 16
    r = wl.poisson_radius_for(262144)
    smp = vmhip.Sampler(W, H, r, 30, 1)
    grid, gw, gh, cs = smp.cells(), smp.grid_w, smp.grid_h, smp.cell_size
    sites = wl.live_sites(grid)
    targets = wl.ellipses(4194304, count=64, sigma=2e-4, seed=1)
    S, T = len(sites), len(targets)
    print(f"case: large instance, {S} sites, {T} targets", flush=True)
    opt = vmhip.Optimizer(W, H, r, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / r, iterations=500)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    so = _slots(grid)
    g = opt.loss_pass(T, want_terms=False)["grad"][so]
    cell = 2.0 * r
    r32 = O.loss_grad(sites, targets, K, W, H, cell, per_target=True)
    assert np.array_equal(g, r32["grad"])
    s64, t64 = sites.astype(np.float64), targets.astype(np.float64)
    r64 = O.loss_grad(s64, t64, K, W, H, cell, real=1, per_target=True)
    flip = (r32["c0"] != r64["c0"]) | (r32["bestj"] != r64["bestj"])
    keep = ~flip & (r32["c0"] >= 0)
    gk32, gk64 = _grad_from_terms(r32, keep, S), _grad_from_terms(r64, keep, S)
    e_keep = np.linalg.norm(gk32 - gk64) / np.linalg.norm(gk64)
    e_all = np.linalg.norm(g - r64["grad"]) / np.linalg.norm(r64["grad"])
    print(f"FP32 vs FP64: {int(flip.sum())} of {T} selections differ; gradient rel err {e_keep:.3e} over the others, "
          f"{e_all:.3e} over all targets", flush=True)
    assert flip.sum() <= 8 and e_keep < 1e-5
    opt64 = vmhip.Optimizer(W, H, r, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 / r, iterations=500, precision=1)
    opt64.set_targets(targets)
    opt64.set_generators(grid, gw, gh, cs)
    g64 = opt64.loss_pass(T, want_terms=False)["grad"][so]
    e64 = np.linalg.norm(g64 - r64["grad"]) / np.linalg.norm(r64["grad"])
    print(f"FP64 mode vs FP64 oracle: gradient rel err {e64:.3e}", flush=True)
    assert e64 < 1e-12
    opt64.close()
    opt.step(3)
    ref, _, _, lh, ch = O.fit(sites, targets, K, cell, O.make_sched(_lr(r), W, H), 3, box=_box(r, gw, gh, cs))
    assert np.array_equal(opt.generators()[so], ref)
    assert ch.min() > 4_000_000


def test_boundary_box_sides_and_empty_occupancy():
    """Every voxel occupied: every cell is inside and the boundary is the domain box (keys with box-side labels
    -1..-4), one loop; no voxel occupied: no boundary.  Both bit-exact against the oracle's assembly."""
    r, K = 0.04, 16
    grid = O.poisson(W, H, r, 30, 5)
    sites = wl.live_sites(grid)
    b = vmhip.Batch()
    b.add(W, H, sites, wl.star(8 * len(sites), seed=5), 1e-3, loss_k=K)
    b.commit()
    b.step(2)
    live = b.sites(0)
    tx = ty = 8
    vox = np.float32(1.0 / 64)
    for name, word in (("full", ~np.uint64(0)), ("empty", np.uint64(0))):
        print(f"case: boundary with {name} occupancy", flush=True)
        masks = np.full(tx * ty, word, np.uint64)
        m = b.extract(0, masks, (tx, ty), vox)
        ref = O.clip(live, W, H, 0.1, masks=masks, tiles=(tx, ty), voxel=vox)
        assert np.array_equal(m["inside"], ref["inside"])
        bd = O.boundary(ref["verts"], ref["labels"], ref["count"], ref["inside"])
        assert np.array_equal(m["vertices"], bd["vertices"]) and np.array_equal(m["edges"], bd["edges"])
        assert m["watertight"] == bd["watertight"] and m["loops"] == bd["loops"]
        if name == "full":
            assert m["inside"].all() and m["watertight"] and m["loops"] == 1 and len(m["edges"]) > 0
        else:
            assert not m["inside"].any() and len(m["edges"]) == 0


def test_loss_pass_matches_committed_fixture():
    """The GPU loss pass on tests/golden/config0_small.npz (r = 0.05, 1,024 star targets, k = 16): k-NN lists and
    per-target terms bit-exact against the stored oracle values, the FP32 gradient within 1e-5 (norm-wise)."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "config0_small.npz"))
    grid, targets = z["grid"], z["targets"]
    gw, gh, cs = O.poisson_dims(W, H, 0.05)
    opt = _optimizer(0.05, 16)
    opt.set_targets(targets)
    opt.set_generators(grid, gw, gh, cs)
    d = opt.loss_pass(len(targets))
    opt.close()
    slot = _slots(grid)
    assert np.array_equal(d["knn"], _map(z["knn"], slot))
    assert np.array_equal(d["loss_terms"], z["loss_t"].astype(np.float64))
    g = d["grad"][slot]
    assert np.linalg.norm(g - z["grad"]) / np.linalg.norm(z["grad"]) < 1e-5


@pytest.mark.parametrize("gap", [1e-3, 1e-6])
def test_gradient_clamp_counted(gap):
    """FP32 gradient sums are fixed point, each term clamped to 2^(51 - shift) first (shift 47 for 64 targets in
    the unit square, so |term| <= 16).  A site pair `gap` apart with 64 targets on a ring of radius 0.03 around it:
    at 1e-3 no term reaches the clamp; at 1e-6 the pair's terms (about 2 d |u_perp| / gap) pass it.  The reported
    count equals the targets whose FP32 terms (the oracle's per-target terms, which the kernel computes bit for bit)
    pass the clamp, the clamped sums equal the oracle's, and against the unclamped FP64 gradient only the pair's
    two sites differ (by more than 1e-5 of the largest other gradient)."""
    r, K = 0.1, 16
    grid = O.poisson(W, H, r, 30, 5)
    gw, gh, cs = O.poisson_dims(W, H, r)
    so = _slots(grid)
    g = grid.copy()
    a, b = 10, 11
    g[so[b]] = g[so[a]] + np.float32([gap, 0.0])
    ang = np.linspace(0.0, 2 * np.pi, 64, endpoint=False) + 0.1
    targets = (g[so[a]] + 0.03 * np.stack([np.cos(ang), np.sin(ang)], 1)).astype(np.float32)
    sites = wl.live_sites(g)
    opt = _optimizer(r, K)
    opt.set_targets(targets)
    opt.set_generators(g, gw, gh, cs)
    d = opt.loss_pass(len(targets))
    sat = opt.grad_saturation()
    opt.close()
    ref = O.loss_grad(sites, targets, K, W, H, 0.1, per_target=True)
    expect = int((np.abs(ref["term_grad"]).max(axis=1) > 2.0 ** (51 - 47)).sum())
    print(f"case: gap {gap}: {sat} clamped targets (oracle terms: {expect})", flush=True)
    assert sat == expect
    assert (sat > 0) == (gap < 1e-5)
    grad = d["grad"][so]
    assert np.array_equal(grad, ref["grad"])
    r64 = O.loss_grad(sites.astype(np.float64), targets.astype(np.float64), K, W, H, 0.1, real=1)
    err = np.abs(grad - r64["grad"]).max(axis=1)
    others = np.setdiff1d(np.arange(len(sites)), [a, b])
    tol = 1e-5 * np.abs(r64["grad"][others]).max()
    assert err[others].max() <= tol
    if sat:  # the clamped terms (16 against hundreds) largely cancel over the ring, but not to within tol
        assert err[[a, b]].min() > tol


@pytest.mark.parametrize("cell_factor,rings", [(3.0, 1), (4.0, 2), (0.5, 2)])
def test_unstaged_chunks_bitexact(cell_factor, rings):
    """Grid shapes whose staged blocks overflow the loss kernel's stage or miss the candidate disc: those chunks
    queue their targets for the exact-selection kernel's ring search.  The batch loss pass (per-target terms, no
    k-NN lists) and five optimizer steps are bit-exact against the oracle."""
    r, K, iters = 0.025, 16, 5
    grid, _, targets = _config0(r)
    sites = wl.live_sites(grid)
    spacing = (1.0 / len(sites)) ** 0.5
    print(f"case: unstaged chunks, cell {cell_factor} spacings, {rings} staged rings", flush=True)
    b = vmhip.Batch()
    b.add(W, H, sites, targets, 1e-3, loss_k=K, grid_cell=cell_factor * spacing, stage_rings=rings)
    b.commit()
    st = b.loss_stats()
    assert st["ring_staged_waves"] + st["ring_global_waves"] > 0
    d = b.loss_pass()
    ref = O.loss_grad(sites, targets, K, W, H, 0.1, per_target=True)
    assert np.array_equal(d["nearest"], ref["c0"])
    assert np.array_equal(d["partner"], ref["bestj"])
    assert np.array_equal(d["loss_terms"], ref["loss_t"].astype(np.float64))
    assert np.array_equal(d["grad"], ref["grad"])
    b.step(iters)
    sc = O.make_sched(1e-3, W, H)
    fit, _, _, lh, ch = O.fit(sites, targets, K, 0.1, sc, iters)
    assert np.array_equal(b.sites(0), fit)
    assert b.status(0)["last_mean_loss"] == pytest.approx(lh[-1] / ch[-1], rel=1e-12)
