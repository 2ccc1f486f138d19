"""The serving loop over a synthetic stream generated inside the library (vm_serve_synthetic): every arrival is
served (drain) or reported as backlog (no drain), pools that start small grow without changing any result, and a
request's result is the same in the stream as when it is regenerated and served alone."""
import numpy as np
import pytest

import vmhip

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


def _check_records(rec, st):
    assert len(rec) == st["arrived"]
    assert np.all(np.diff(rec["arrival_s"]) > 0)
    adm = rec["admit_s"] >= 0
    assert adm.sum() == st["admitted"]
    done = rec["done_s"] >= 0
    assert done.sum() == st["completed"]
    assert np.all(rec["admit_s"][adm] >= rec["arrival_s"][adm])
    assert np.all(rec["done_s"][done] >= rec["admit_s"][done])
    # admission in arrival order: the admitted requests are a prefix of the arrivals
    assert np.all(adm[:st["admitted"]]) and not np.any(adm[st["admitted"]:])


def _alone(stream, i, iters, precision=0):
    req, _ = vmhip.synth_request(stream, i)
    res, _ = vmhip.serve([req], [0.0], max_iterations=iters, tol=0.0, poll_every=5, precision=precision)
    return req, res[0]


@pytest.mark.parametrize("initial_pool_bytes", [0, 2e6], ids=["default-pools", "growing-pools"])
def test_stream_drains_and_matches_alone(initial_pool_bytes):
    iters = 40
    stream = vmhip.serve_stream_desc(rate=300.0, duration_s=0.2, seed=11, min_sites=512, max_sites=8192, drain=True)
    print(f"case: stream rate 300/s 0.2 s, sites 512-8192, initial pools {initial_pool_bytes}", flush=True)
    rec, st = vmhip.serve_synthetic(stream, max_iterations=iters, tol=0.0, poll_every=5,
                                    initial_pool_bytes=initial_pool_bytes)
    _check_records(rec, st)
    assert st["arrived"] >= 30 and st["completed"] == st["arrived"]
    assert np.all(rec["iterations"] == iters)
    assert st["max_resident_seen"] > 1
    if initial_pool_bytes:
        assert st["pool_grows"] >= 3  # the pools started far too small for the stream
    sizes = rec["num_sites"]
    for i in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), len(rec) // 2}):
        req, alone = _alone(stream, i, iters)
        assert len(req["sites"]) == rec["num_sites"][i] and len(req["targets"]) == rec["num_targets"][i]
        assert alone["iterations"] == iters
        # the loss sum is an FP64 atomic reduction (order-dependent in its last bits)
        assert alone["final_mean_loss"] == pytest.approx(rec["final_mean_loss"][i], rel=1e-12), f"request {i}"


def test_stream_overload_reports_backlog():
    """Arrivals far beyond what a 300k-site resident cap can serve in 0.4 s: the loop ends at the window, the
    unadmitted arrivals are backlog (never generated), and nothing resident passed the cap beyond one request."""
    stream = vmhip.serve_stream_desc(rate=2000.0, duration_s=0.4, seed=12, min_sites=512, max_sites=131072,
                                     drain=False)
    print("case: overload stream rate 2000/s 0.4 s, cap 300k sites", flush=True)
    rec, st = vmhip.serve_synthetic(stream, max_iterations=300, tol=1e-6, poll_every=10, max_resident_sites=3e5)
    _check_records(rec, st)
    assert st["arrived"] > 600
    assert st["admitted"] < st["arrived"] and st["oldest_wait_s"] > 0.1
    assert st["completed"] <= st["admitted"]
    assert st["wall_s"] < 0.4 + 2.0
    # resident sites at any admission stayed under the cap (one request may exceed it alone)
    adm = np.nonzero(rec["admit_s"] >= 0)[0]
    events = [(rec["admit_s"][i], rec["num_sites"][i]) for i in adm]
    events += [(rec["done_s"][i], -rec["num_sites"][i]) for i in adm if rec["done_s"][i] >= 0]
    events.sort(key=lambda e: (e[0], e[1]))
    level, peak = 0, 0
    for _, ds in events:
        level += ds
        peak = max(peak, level)
    assert peak <= 3e5 + rec["num_sites"][adm].max()


def test_stream_fp64_matches_alone():
    iters = 20
    stream = vmhip.serve_stream_desc(rate=200.0, duration_s=0.1, seed=13, min_sites=512, max_sites=4096, drain=True)
    print("case: FP64 stream rate 200/s 0.1 s", flush=True)
    rec, st = vmhip.serve_synthetic(stream, max_iterations=iters, tol=0.0, poll_every=5, precision=1)
    _check_records(rec, st)
    assert st["completed"] == st["arrived"] >= 8
    i = int(np.argmax(rec["num_sites"]))
    _, alone = _alone(stream, i, iters, precision=1)
    assert alone["final_mean_loss"] == pytest.approx(rec["final_mean_loss"][i], rel=1e-12)
