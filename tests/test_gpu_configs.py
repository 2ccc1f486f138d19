"""GPU parity at the workloads' full sizes (the mixed batch, the precision stress, a serving sample).

Bars as in test_gpu_parity.py: bit-exact against the oracle's FP32 semantics (k-NN lists, selections, loss
terms, fixed-point gradient sums, positions), FP64 mode within 1e-12 of the FP64 oracle, clipped-cell
topology with zero mismatches.
"""
import numpy as np
import pytest

import oracle as O

vmhip = pytest.importorskip("vmhip")
from vmhip import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

W = H = 1.0


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


def _mixed_batch(n=512, seed=2):
    radii, seeds = wl.batch_plan(n, seed)
    grids, _, _ = vmhip.poisson_batch(np.ones(n), np.ones(n), radii, seeds)
    return radii, seeds, grids


def test_mixed_batch_full_size_bitexact():
    """The 512-instance mixed batch at full size: sites log-uniform in [256, 65,536] from the batched Poisson
    sampler (three instances checked against the oracle's sampler), 8 targets per site on each instance's random
    polygon, k from the tuner's choice for the size class.  Two batched iterations leave every instance's
    positions bit-identical to the oracle's two-step fit with the same k, grid cell and schedule; every instance
    counts two steps, and the loss kernel settles nearly every target on its staged path."""
    radii, seeds, grids = _mixed_batch()
    n = len(grids)
    for i in (0, int(np.argmin(radii)), int(np.argmax(radii))):
        assert np.array_equal(grids[i], O.poisson(W, H, float(radii[i]), 30, int(seeds[i])))
    insts = []
    b = vmhip.Batch()
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        tg = wl.batch_targets(sites, seeds[i])
        insts.append((sites, tg))
        b.add(W, H, sites, tg, 1e-3, loss_k=0)
    S = sum(len(s) for s, _ in insts)
    T = sum(len(t) for _, t in insts)
    print(f"case: mixed batch, {n} instances, {S} sites, {T} targets", flush=True)
    assert 5_500_000 < S < 7_500_000
    b.commit()
    st = b.loss_stats()
    assert st["ring_global_waves"] == 0
    assert st["fast_waves"] > 0.99 * (st["fast_waves"] + st["ring_staged_waves"])
    assert st["exact_fallback_targets"] < 1e-3 * T
    iters = 2
    b.step(iters)
    sc = O.make_sched(1e-3, W, H)
    ks = set()
    for i, (sites, tg) in enumerate(insts):
        ch = b.choice(i)
        ks.add(ch["loss_k"])
        ref, _, _, lh, ch_ = O.fit(sites, tg, ch["loss_k"], 2.0 / np.sqrt(len(sites)), sc, iters)
        got = b.sites(i)
        if not np.array_equal(got, ref):
            raise AssertionError(f"instance {i} ({len(sites)} sites, k={ch['loss_k']}) differs from the oracle")
        s = b.status(i)
        assert s["iteration"] == iters
        assert s["last_mean_loss"] == pytest.approx(lh[-1] / ch_[-1], rel=1e-12)
    print(f"k chosen: {sorted(ks)}", flush=True)


@pytest.mark.parametrize("precision", [0, 1], ids=["fp32", "fp64"])
def test_precision_stress_full_size(precision):
    """The precision stress at full size: 65,536 sites on a 256 x 256 lattice jittered by +-1e-7 (many nearly
    concurrent bisector triples), 524,288 targets on a circle of radius 0.3 (sigma 1e-4), k = 16.
    FP32 mode: k-NN lists, nearest sites, bisector partners, loss terms and the fixed-point gradient bit-exact
    against the FP32 oracle, and five optimizer steps bit-exact.  FP64 mode: selections and terms equal to the
    FP64 oracle's, gradient within 1e-12 and positions after five steps within 1e-12 before their rounding to
    FP32 at the C ABI.  Both: the clipped cells (FP64
    predicates) equal the FP64 oracle's in every label and vertex: zero topology mismatches."""
    K, iters = 16, 5
    sites = wl.lattice(256, 1e-7, seed=3)
    tg = wl.circle(524288, radius=0.3, sigma=1e-4, seed=3)
    print(f"case: precision stress, precision {precision}", flush=True)
    opt = vmhip.Optimizer(W, H, 1 / 256, loss_k=K, sample_fraction=1.0, lr_per_spacing=1e-3 * 256,
                          iterations=iters, precision=precision)
    opt.set_targets(tg)
    opt.set_generators(sites, 256, 256, 1 / 256)
    d = opt.loss_pass(len(tg))
    cell = 2 / 256
    real = precision
    s_in = sites.astype(np.float64) if real else sites
    t_in = tg.astype(np.float64) if real else tg
    ref = O.loss_grad(s_in, t_in, K, W, H, cell, real=real, per_target=True)
    assert np.array_equal(d["nearest"], ref["c0"])
    assert np.array_equal(d["partner"], ref["bestj"])
    assert np.array_equal(d["loss_terms"], ref["loss_t"].astype(np.float64))
    assert d["count"] == ref["count"] == len(tg)
    if precision == 0:
        ridx, _ = O.knn(sites, tg, K, W, H, cell)
        assert np.array_equal(d["knn"], ridx)
        assert np.array_equal(d["grad"], ref["grad"])
    else:
        assert np.linalg.norm(d["grad"] - ref["grad"]) / np.linalg.norm(ref["grad"]) < 1e-12
    mesh = opt.extract()
    cref = O.clip(sites.astype(np.float64), W, H, cell, real=1)
    cnt = mesh["cell_start"][1:] - mesh["cell_start"][:-1]
    assert np.array_equal(cnt, cref["count"])
    mism = 0
    for i in range(len(sites)):
        b0, n = mesh["cell_start"][i], cnt[i]
        mism += not (np.array_equal(mesh["cell_labels"][b0:b0 + n], cref["labels"][i, :n]) and
                     np.array_equal(mesh["cell_verts"][b0:b0 + n], cref["verts"][i, :n]))
    print(f"topology mismatches {mism}, unsafe {mesh['unsafe_cells']}, overflow {mesh['overflow_cells']}",
          flush=True)
    assert mism == 0 and mesh["overflow_cells"] == cref["overflow"] == 0
    assert mesh["unsafe_cells"] == cref["unsafe"]
    opt.step(iters)
    lr = float(np.float32(np.float32(1e-3 * 256) * np.float32(1 / 256)))
    f = np.float32
    m = f(2) * f(1 / 256)
    box = (float(-m), float(-m), float(f(256) * f(1 / 256) + m), float(f(256) * f(1 / 256) + m))
    pos, _, _, _, _ = O.fit(s_in, t_in, K, cell, O.make_sched(lr, W, H), iters, box=box, real=real)
    got = opt.generators()  # FP32 at the C ABI in both modes
    if precision == 0:
        assert np.array_equal(got, pos)
    else:
        # Within 1e-12 of the FP64 oracle before the final rounding to FP32 (half an ulp below 1: 2^-25).
        assert np.abs(got.astype(np.float64) - pos).max() <= 2.0 ** -25 + 1e-12


def _serve_requests(n, seed):
    """Requests drawn from the serving workload's law: site counts log-uniform in [512, 131,072] (batched Poisson
    sampler), 8 targets per site on a random polygon, lr 1e-3, k = 16."""
    sizes = wl.log_uniform_sizes(n, 512, 131072, seed)
    radii = np.array([wl.poisson_radius_for(int(s)) for s in sizes], np.float32)
    seeds = (4 * (1 << 20) + 1000 * seed + np.arange(n)).astype(np.uint32)
    grids, _, _ = vmhip.poisson_batch(np.ones(n), np.ones(n), radii, seeds)
    reqs = []
    for i, g in enumerate(grids):
        sites = wl.live_sites(g)
        reqs.append(dict(width=1.0, height=1.0, sites=sites, targets=wl.batch_targets(sites, seeds[i]), lr=1e-3,
                         loss_k=16))
    return reqs


def _request_bytes(req):
    """Resident bytes the server charges for a request (FP32): per site 4 positions-sized arrays, two int64
    gradient words, three ints and a flag; per target position, slot and queue entry; per hash cell two ints."""
    S, T = len(req["sites"]), len(req["targets"])
    r64 = lambda v: (max(v, 1) + 63) // 64 * 64  # noqa: E731
    cell = np.float32(max(1.5, 0.5 * ((16 / np.pi) ** 0.5 + 0.8)) * (1.0 / S) ** 0.5)
    C = int(np.ceil(np.float32(1.0) / cell)) ** 2 + 1
    return r64(S) * (4 * 8 + 2 * 8 + 12 + 1) + r64(T) * (8 + 4 + 8) + r64(C) * 8


def test_serving_sample_full_size_law():
    """Eight requests from the serving workload's size law (512-131,072 sites) arriving over 6 ms under a tight
    HBM budget (the largest request plus a quarter): the budget holds at every admission, requests overlap on the
    device, and each result (20 iterations) is bit-identical alone and in the mix and to the oracle's fit."""
    reqs = _serve_requests(8, seed=7)
    sizes = [len(r["sites"]) for r in reqs]
    print(f"case: serving sample, sites {sizes}", flush=True)
    need = [_request_bytes(r) for r in reqs]
    budget = 1.25 * max(need)
    assert sum(need) > 2 * budget  # the budget forces queueing
    arrivals = np.array([0.0, 0.0, 0.001, 0.002, 0.003, 0.004, 0.005, 0.006])
    iters = 20
    mix, st = vmhip.serve(reqs, arrivals, hbm_budget_bytes=budget, max_iterations=iters, tol=0.0, poll_every=5,
                          want_sites=True)
    assert st["completed"] == len(reqs)
    assert st["peak_bytes"] <= budget
    assert st["max_resident_seen"] >= 2
    assert all(r["iterations"] == iters for r in mix)
    sc = O.make_sched(1e-3, 1.0, 1.0)
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes)), 3):
        alone, _ = vmhip.serve([reqs[i]], [0.0], hbm_budget_bytes=budget, max_iterations=iters, tol=0.0,
                               poll_every=5, want_sites=True)
        assert np.array_equal(alone[0]["sites"], mix[i]["sites"])
        # the loss sum is an FP64 atomic reduction (order-dependent in its last bits); positions are exact
        assert alone[0]["final_mean_loss"] == pytest.approx(mix[i]["final_mean_loss"], rel=1e-12)
        ref, _, _, _, _ = O.fit(reqs[i]["sites"], reqs[i]["targets"], 16, 2.0 / np.sqrt(sizes[i]), sc, iters)
        assert np.array_equal(mix[i]["sites"], ref), f"request {i} ({sizes[i]} sites) differs from the oracle"
