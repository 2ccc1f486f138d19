"""vm_poisson_batch_noalloc (the batched Poisson sampler over the library's fixed device array, no allocation)
against vm_poisson_batch through the C ABI: identical grids on two calls in a row, the second with a smaller
input, and sizes beyond the array rejected with VM_EINVAL before any GPU work."""
import ctypes as C

import numpy as np
import pytest

vmhip = pytest.importorskip("vmhip")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")


def _run(radii, seeds, noalloc):
    n = len(radii)
    grids, dims, ms = vmhip.poisson_batch(np.ones(n), np.ones(n), radii, seeds, noalloc=noalloc)
    return grids, dims


def test_noalloc_matches_allocating_entry_point_twice():
    # First call: 40 instances (about 150 k grid cells); second: 7 smaller ones, so the array still holds
    # the first call's grids beyond the second call's (and the second call's cells are all rewritten).
    for step, (radii, seeds) in enumerate([
            (np.linspace(0.004, 0.05, 40).astype(np.float32), np.arange(40, dtype=np.uint32) + 11),
            (np.linspace(0.02, 0.08, 7).astype(np.float32), np.arange(7, dtype=np.uint32) + 900)]):
        print(f"case: noalloc call {step}, {len(radii)} instances", flush=True)
        ref, dims = _run(radii, seeds, noalloc=False)
        got, dims2 = _run(radii, seeds, noalloc=True)
        assert dims == dims2
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        assert sum(len(g) for g in got) > 1000


def test_noalloc_rejects_oversize_input():
    # 3 M cells (r = 0.0008 in the unit square) exceed the 16 MiB array: VM_EINVAL with a message.
    lib = vmhip.lib()
    r = np.array([0.0008], np.float32)
    w = np.ones(1, np.float32)
    s = np.ones(1, np.uint32)
    off = np.zeros(1, np.int64)
    cells = np.zeros((1, 2), np.float32)  # never written: the call fails before any GPU work
    ms = C.c_float()
    rc = lib.vm_poisson_batch_noalloc(1, w.ctypes.data, w.ctypes.data, r.ctypes.data, s.ctypes.data, 30,
                                      off.ctypes.data, cells.ctypes.data, C.byref(ms))
    assert rc == 1
    assert b"device scratch" in lib.vm_last_error()
    # and the entry point still works afterwards
    got, _ = _run(np.array([0.03], np.float32), np.array([5], np.uint32), noalloc=True)
    ref, _ = _run(np.array([0.03], np.float32), np.array([5], np.uint32), noalloc=False)
    assert np.array_equal(got[0], ref[0])
