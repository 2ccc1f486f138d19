"""CPU: the benchmark's gradient gate (bench.py) -- the FP64 gradient re-evaluated with the FP32 run's choices.

Pinned against the oracle: for a fit whose FP32 and FP64 selections are forced to differ on some targets, the gate's
corrected FP64 gradient equals the FP64 sum of every target's term taken at the FP32 selection (computed here from
the oracle's FP64 per-target term gradients where the selections agree and from the reference formula elsewhere)."""
import numpy as np

import bench
import oracle as O
from vmhip import workloads as wl


def test_gate_reevaluates_flipped_targets_with_fp32_choices():
    g = O.poisson(1.0, 1.0, 0.04, 30, 3)
    sites = wl.live_sites(g)
    tg = wl.star(4000, seed=4)
    K = 16
    r32 = O.loss_grad(sites, tg, K, 1, 1, 0.1, per_target=True)
    r64 = O.loss_grad(sites.astype(np.float64), tg.astype(np.float64), K, 1, 1, 0.1, real=1, per_target=True)
    # Plant flips: pretend the FP64 run chose another partner for 25 targets (a valid neighbour pair).
    rng = np.random.default_rng(0)
    p64 = dict(grad=r64["grad"].copy(), nearest=r64["c0"].copy(), partner=r64["bestj"].copy())
    used = np.nonzero(r64["c0"] >= 0)[0]
    flip = rng.choice(used, 25, replace=False)
    for t in flip:
        other = (p64["partner"][t] + 1) % len(sites)
        if other == p64["nearest"][t]:
            other = (other + 1) % len(sites)
        g0, gj = bench.term_grads64(sites, tg, np.array([t]), p64["nearest"][t:t + 1], p64["partner"][t:t + 1])
        p64["grad"][p64["nearest"][t]] -= g0[0]
        p64["grad"][p64["partner"][t]] -= gj[0]
        g0, gj = bench.term_grads64(sites, tg, np.array([t]), p64["nearest"][t:t + 1], np.array([other]))
        p64["grad"][p64["nearest"][t]] += g0[0]
        p64["grad"][other] += gj[0]
        p64["partner"][t] = other
    p32 = dict(grad=r32["grad"], nearest=r32["c0"], partner=r32["bestj"])
    gate = bench.gradient_gate([sites], [tg], p32, p64)
    assert gate["selection_flips"] >= 25
    # The teacher-forced FP64 gradient: every used target's FP64 term at the FP32 choice.
    ok = (r32["c0"] >= 0) & (r32["bestj"] >= 0)
    t_idx = np.nonzero(ok)[0]
    g0, gj = bench.term_grads64(sites, tg, t_idx, r32["c0"][ok], r32["bestj"][ok])
    ref = np.zeros((len(sites), 2))
    np.add.at(ref, r32["c0"][ok], g0)
    np.add.at(ref, r32["bestj"][ok], gj)
    # the oracle's own FP64 gradient agrees with the formula where no flip was planted
    assert np.allclose(ref, p32["grad"], rtol=1e-4, atol=1e-9)
    err = np.linalg.norm(p32["grad"] - ref) / np.linalg.norm(ref)
    assert abs(gate["grad_rel_norm_err"] - err) <= 1e-9 * max(err, 1e-12) + 1e-15
    assert gate["passed"] == (err <= bench.GRAD_GATE)
