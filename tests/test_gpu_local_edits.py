"""Local edits of the plain optimizer without a re-layout: UpdateTargets retires the dirty slots' sorted entries and
lays their new targets out as chunks of their own, UpdateGenerators copies the slot rectangle and bins again,
BeginSchedule rewrites the record.  The results equal those of the same sequence with every edit re-laid out
(VM_EDIT_FULL=1), bit for bit, and no edit re-captures the iteration graph."""
import numpy as np
import pytest

import vmhip
from vmhip import workloads as wl

pytestmark = pytest.mark.gpu

W = H = 1.0


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


STAMPS = [((0.5, 0.9), 0.05, True), ((0.5, 0.32), 0.04, False), ((0.2, 0.55), 0.06, True),
          ((0.8, 0.55), 0.03, False), ((0.62, 0.2), 0.05, True), ((0.35, 0.75), 0.04, True)]


def _sequence(full, monkeypatch, r=0.025, K=12, iters=60, local=30):
    monkeypatch.setenv("VM_EDIT_FULL", "1" if full else "0")
    canvas = vmhip.Canvas(W, H, 1.0 / 1024)
    canvas.rasterize(wl.star_outline().astype(np.float32))
    sampler = vmhip.Sampler(W, H, r, 30, 1)
    gw, gh, cs = sampler.grid_w, sampler.grid_h, sampler.cell_size
    slots = canvas.target_slots()
    opt = vmhip.Optimizer(W, H, r, loss_k=K, sample_fraction=0.2, lr_per_spacing=1e-3 / r, iterations=iters)
    # Room for the canvas to grow: every UpdateTargets below stays within the capacity.
    opt.set_targets(np.concatenate([slots, np.full((len(slots) // 2, 2), -1.0, np.float32)]))
    opt.set_generators(sampler.cells(), gw, gh, cs)
    opt.step(iters)
    graphs = opt.graph_stats()
    out = [(opt.generators(), opt.evaluate_loss())]
    for c, brush, add in STAMPS:
        _, dirty = canvas.stamp_disc(c, brush, add)
        opt.update_targets(dirty, canvas.target_slots())
        margin = brush + 2 * r
        rect = (max(int(np.floor((c[0] - margin) / cs)), 0), max(int(np.floor((c[1] - margin) / cs)), 0),
                min(int(np.ceil((c[0] + margin) / cs)), gw), min(int(np.ceil((c[1] + margin) / cs)), gh))
        sampler.resample(*rect)
        opt.update_generators(sampler.cells(), *rect)
        opt.begin_schedule(local, rect)
        opt.step(local)
        out.append((opt.generators(), opt.evaluate_loss()))
    stats = opt.graph_stats()
    opt.close()
    return out, graphs, stats


def test_incremental_edits_match_full_relayout(monkeypatch):
    print("case: 6 stamps, incremental edits", flush=True)
    inc, g0, g1 = _sequence(False, monkeypatch)
    print("case: 6 stamps, every edit re-laid out", flush=True)
    full, _, _ = _sequence(True, monkeypatch)
    assert g1 == g0  # no edit changed the iteration's launches
    for k, ((pa, la), (pb, lb)) in enumerate(zip(inc, full)):
        assert np.array_equal(pa, pb), f"positions differ after stamp {k}"
        # the mean loss is an FP64 reduction over a different chunking (last bits)
        assert la == pytest.approx(lb, rel=1e-12)
