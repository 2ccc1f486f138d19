"""The reference demo's call sequences (Editor::Reset, Editor::Stamp) through the
header-only C++ drop-in API, as a compiled program (tests/cpp/dropin_demo.cpp)."""
import json
import os
import subprocess

import pytest

import vmhip

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_demo")


def test_cpp_dropin_reset_and_stamp():
    if vmhip.device_count() == 0:
        pytest.skip("no HIP device")
    assert os.path.exists(BIN), "tests/cpp/dropin_demo is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert r.returncode == 0 and out["ok"], out
    assert out["watertight"] == [1, 1] and out["loops"][0] == 1
    assert out["overflow"] == [0, 0]
    assert out["loss1"] < out["loss0"]
