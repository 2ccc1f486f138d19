"""GPU: the iteration graphs after a new layout.

`step` replays two captured graphs (one iteration, eight iterations).  After a
commit the launches are captured again and the existing execs take them in
place when the graph keeps its shape, or are instantiated again when it does
not.  `set_generators` zeroes the moments and restarts the schedule, so an
optimizer that ran before must then give the same positions, bit for bit, as a
fresh one given the same generators.  Going from about 1,100 sites (one scan
tile, scatter folded into the scan) to about 80,000 (several tiles, a scatter
launch) changes the graph's shape; a second grid of that size keeps it.  The
optimizer's graph counters show which path each re-layout took.
"""
import numpy as np
import pytest

vmhip = pytest.importorskip("vmhip")
from vmhip import workloads as wl  # noqa: E402

pytestmark = pytest.mark.gpu

W = H = 1.0
R_SMALL = 0.025


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


def _grid(r, seed):
    s = vmhip.Sampler(W, H, r, 30, seed)
    return s.cells(), s.grid_w, s.grid_h, s.cell_size


def _optimizer(targets):
    opt = vmhip.Optimizer(W, H, R_SMALL, loss_k=16, sample_fraction=1.0, lr_per_spacing=1e-3 / R_SMALL,
                          iterations=200)
    opt.set_targets(targets)
    return opt


def test_graph_update_after_relayout_matches_fresh_optimizer():
    targets = wl.star(8192, seed=31)
    small = _grid(R_SMALL, 31)
    r_large = wl.poisson_radius_for(80000)
    large = [_grid(r_large, 32), _grid(r_large, 33)]
    reused = _optimizer(targets)
    reused.set_generators(*small)
    reused.step(9)  # one eight-iteration replay and one single-iteration replay
    assert reused.graph_stats() == dict(updates=0, instantiations=2)
    # new shape: both execs instantiated again; same shape: both taken in place
    expect = [dict(updates=0, instantiations=4), dict(updates=2, instantiations=4)]
    for n, g in enumerate(large):
        print(f"case: re-layout {n} ({'new shape' if n == 0 else 'same shape'}) against a fresh optimizer", flush=True)
        reused.set_generators(*g)
        reused.step(9)
        assert reused.graph_stats() == expect[n], f"re-layout {n}: {reused.graph_stats()}"
        fresh = _optimizer(targets)
        fresh.set_generators(*g)
        fresh.step(9)
        assert np.array_equal(reused.generators(), fresh.generators()), f"re-layout {n}"
        fresh.close()
