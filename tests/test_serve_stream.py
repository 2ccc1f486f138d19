"""The synthetic serving stream's generator (host code of the library, no device): each request is a pure function
of (seed, index), and the draws follow the stated laws -- Poisson arrivals, log-uniform site counts, targets on a
random polygon with Gaussian jitter."""
import math

import numpy as np
import pytest
from scipy import stats

import vmhip


def _stream(**kw):
    base = dict(rate=2000.0, duration_s=60.0, seed=4, min_sites=512, max_sites=131072, targets_per_site=8,
                sigma=5e-4, lr=1e-3, loss_k=16)
    base.update(kw)
    return vmhip.serve_stream_desc(**base)


def test_plan_is_a_function_of_seed_and_index():
    st = _stream()
    a = [vmhip.synth_plan(st, i) for i in (0, 1, 57)]
    b = [vmhip.synth_plan(st, i) for i in (0, 1, 57)]
    assert a == b
    other = [vmhip.synth_plan(_stream(seed=5), i) for i in (0, 1, 57)]
    assert all(x[1:] != y[1:] for x, y in zip(a, other))


def test_arrivals_are_poisson_and_sizes_log_uniform():
    rate, n = 2000.0, 4000
    st = _stream(rate=rate)
    plans = [vmhip.synth_plan(st, i) for i in range(0, n, 1)]
    t = np.array([p[0] for p in plans])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    # exponential gaps with mean 1/rate
    assert stats.kstest(gaps * rate, "expon").pvalue > 1e-3
    s = np.array([p[1] for p in plans])
    assert s.min() >= 512 and s.max() <= 131072
    u = (np.log(s) - math.log(512)) / (math.log(131072) - math.log(512))
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    r = np.array([p[2] for p in plans], np.float32)
    assert np.array_equal(r, np.sqrt(0.668 / s).astype(np.float32))
    # arrival of request i is the sum of the first i + 1 gaps
    assert vmhip.synth_plan(st, 10)[0] == pytest.approx(t[10], rel=1e-15)


def test_targets_lie_on_the_polygon_with_gaussian_jitter():
    sigma = 5e-4
    st, st0 = _stream(sigma=sigma), _stream(sigma=0.0)
    for i in (3, 11):
        n_sites = 3000
        a = vmhip.synth_targets(st, i, n_sites)
        b = vmhip.synth_targets(st, i, n_sites)
        assert a.shape == (8 * n_sites, 2) and np.array_equal(a, b)
        clean = vmhip.synth_targets(st0, i, n_sites).astype(np.float64)
        # the same stream positions without jitter: the difference is sigma * N(0, 1) per axis
        z = (a.astype(np.float64) - clean) / sigma
        assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03
        assert stats.kstest(z.ravel(), "norm").pvalue > 1e-4
        # un-jittered points lie within the polygon's radius band around (0.5, 0.5)
        rad = np.hypot(clean[:, 0] - 0.5, clean[:, 1] - 0.5)
        assert rad.max() <= 0.45 + 1e-6
        # and on its edges: sorted by angle, consecutive points turn by no more than the vertices do
        ang = np.sort(np.arctan2(clean[:, 1] - 0.5, clean[:, 0] - 0.5))
        assert np.diff(ang).max() < 2 * math.pi / 5
    assert not np.array_equal(vmhip.synth_targets(st, 3, 100), vmhip.synth_targets(st, 4, 100))


def test_invalid_streams_are_rejected():
    with pytest.raises(vmhip.VmError):
        vmhip.synth_plan(_stream(rate=0.0), 0)
    with pytest.raises(vmhip.VmError):
        vmhip.synth_plan(_stream(min_sites=600, max_sites=500), 0)
    # a target buffer one point short of targets_per_site * num_sites
    buf = np.zeros((79, 2), np.float32)
    st = _stream()
    rc = vmhip.lib().vm_serve_synth_targets(vmhip.C.byref(st), 0, 10, buf.ctypes.data, 79)
    assert rc != 0 and b"target buffer" in vmhip.lib().vm_last_error()
