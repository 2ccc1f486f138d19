"""C ABI of libvoromesh_hip.so (no GPU needed): the library loads, exports every
entry point include/voromesh_hip.h declares, and rejects invalid arguments with a
status code and a message before touching the device."""
import ctypes as C
import os
import re
import subprocess

import pytest

import vmhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "voromesh_hip.h")


def _declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"VM_API\s+[\w\s\*]+?\b(vm_\w+)\s*\(", src)))


def test_header_declares_the_wrapper_list():
    assert _declared() == sorted(vmhip.EXPORTS)


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", vmhip.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (vm_\w+)", out.stdout))
    missing = [s for s in _declared() if s not in exported]
    assert not missing, missing
    L = vmhip.lib()
    for s in _declared():
        assert getattr(L, s) is not None


def test_version_and_defaults():
    assert b"gfx950" in vmhip.lib().vm_version()
    p = vmhip.default_params()
    assert (p.iterations, p.local_iterations, p.loss_k, p.seed) == (400, 100, 12, 1)
    assert abs(p.lr_per_spacing - 0.09) < 1e-7 and abs(p.sample_fraction - 0.2) < 1e-7
    assert list(p.milestones) == [80, 150, 200, 250] and abs(p.gamma - 0.5) < 1e-7


def test_invalid_arguments_are_rejected_before_the_device():
    L = vmhip.lib()
    h = C.c_void_p()
    p = vmhip.default_params()
    assert L.vm_create(0.0, 1.0, 0.1, C.byref(p), C.byref(h)) == 1
    assert b"positive" in L.vm_last_error()
    assert L.vm_create(1.0, 1.0, 0.1, C.byref(p), None) == 1
    p.loss_k = 7
    assert L.vm_create(1.0, 1.0, 0.1, C.byref(p), C.byref(h)) == 1
    assert L.vm_set_targets(None, None, 0) == 1
    assert L.vm_step(None, 1, None) == 1
    assert L.vm_batch_create(3, C.byref(h)) == 1
    assert L.vm_sampler_create(1.0, 1.0, -0.1, 30, 1, C.byref(h)) == 1
    gw, gh, cs = C.c_int(), C.c_int(), C.c_float()
    assert L.vm_poisson_grid_dims(1.0, 1.0, 0.025, C.byref(gw), C.byref(gh), C.byref(cs)) == 0
    assert (gw.value, gh.value) == (57, 57)


def test_device_count_does_not_fail():
    assert vmhip.device_count() >= 0


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; asked for more it exits non-zero with a message and no JSON line."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr) and "{" not in r.stdout
