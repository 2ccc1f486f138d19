"""Continuous-batching serving loop (vm_serve): requests arriving while others
run are admitted as slots and budget free up, and each result is bit-identical
alone and in the mix (order-independent fixed-point gradient sums)."""
import numpy as np
import pytest

import oracle as O
import vmhip
from vmhip import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


def _requests(n, seed=0):
    sizes = wl.log_uniform_sizes(n, 256, 4096, seed)
    reqs = []
    for i, s in enumerate(sizes):
        g = O.poisson(1.0, 1.0, wl.poisson_radius_for(int(s)), 30, 100 + i)
        sites = wl.live_sites(g)
        reqs.append(dict(width=1.0, height=1.0, sites=sites, targets=wl.polygon(8 * len(sites), seed=200 + i),
                         lr=1e-3, loss_k=16))
    return reqs


def test_serve_result_same_alone_and_in_mix():
    reqs = _requests(12)
    # half arrive together, the rest while the first ones are still running
    arrivals = np.concatenate([np.zeros(6), np.linspace(0.001, 0.003, 6)])
    mix, st = vmhip.serve(reqs, arrivals, max_iterations=60, tol=0.0, poll_every=5, want_sites=True)
    assert st["completed"] == len(reqs)
    assert st["max_resident_seen"] > 1  # requests overlapped on the device
    for i in (0, 5, 11):
        alone, _ = vmhip.serve([reqs[i]], [0.0], max_iterations=60, tol=0.0, poll_every=5, want_sites=True)
        assert alone[0]["iterations"] == mix[i]["iterations"] == 60
        assert np.array_equal(alone[0]["sites"], mix[i]["sites"])
        assert alone[0]["final_mean_loss"] == pytest.approx(mix[i]["final_mean_loss"], rel=1e-12)
        # and equal to the oracle's 60-step fit
        sc = O.make_sched(1e-3, 1.0, 1.0)
        ref, _, _, _, _ = O.fit(reqs[i]["sites"], reqs[i]["targets"], 16, 0.1, sc, 60)
        assert np.array_equal(mix[i]["sites"], ref)


def test_serve_admission_limited_by_slots_and_budget():
    reqs = _requests(10, seed=3)
    arrivals = np.zeros(len(reqs))
    res, st = vmhip.serve(reqs, arrivals, max_resident=3, max_iterations=20, tol=0.0, poll_every=4)
    assert st["completed"] == 10 and st["max_resident_seen"] <= 3
    lat = np.array([r["latency_s"] for r in res])
    q = np.array([r["queue_s"] for r in res])
    assert np.all(lat >= q) and np.all(q >= 0)
    assert q[-1] > q[0]  # later requests waited for a slot
    # four copies of one request under a budget of 1.5 requests: one resident at a time
    one = reqs[0]
    S, T = len(one["sites"]), len(one["targets"])
    r64 = lambda n: (n + 63) // 64 * 64  # noqa: E731
    cell = np.float32(max(1.5, 0.5 * ((16 / np.pi) ** 0.5 + 0.8)) * (1.0 / S) ** 0.5)
    C = int(np.ceil(np.float32(1.0) / cell)) ** 2 + 1
    est = r64(S) * (4 * 8 + 2 * 8 + 12 + 1) + r64(T) * 12 + r64(C) * 8
    budget = 1.5 * est
    res2, st2 = vmhip.serve([one] * 4, np.zeros(4), hbm_budget_bytes=budget, max_iterations=10, tol=0.0,
                            poll_every=2)
    assert st2["completed"] == 4 and st2["max_resident_seen"] == 1 and st2["peak_bytes"] <= budget


def test_serve_convergence_stops_early():
    reqs = _requests(3, seed=5)
    res, _ = vmhip.serve(reqs, [0, 0, 0], max_iterations=300, tol=1e-2, poll_every=1)
    assert all(r["iterations"] < 300 for r in res)
