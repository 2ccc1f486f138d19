"""Headless end-to-end run of the reference demo's scripted edit sequence
(apps/voromesh_demo.cpp:526-570, `--shot`) through the drop-in C++ API
(tests/cpp/shot_demo): heart on a 1024^2 canvas, Poisson r = 24 generators,
global fit, then the three scripted drags with 100 local iterations per stamp.
The demo's own pass criterion is a watertight final mesh; the accuracy bounds
below are the demo's diagnostics (surface-to-mesh distance in voxels)."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHOT = os.path.join(ROOT, "tests", "cpp", "shot_demo")


@pytest.mark.gpu
def test_headless_shot_is_watertight():
    assert os.path.exists(SHOT), "tests/cpp/shot_demo is not built"
    print("case: shot_demo 24 1024 100 --json", flush=True)
    r = subprocess.run([SHOT, "24", "1024", "100", "--json"], capture_output=True, text=True, timeout=110)
    print(r.stdout[-4000:])
    assert r.returncode == 0, r.stderr[-2000:]
    s = json.loads(r.stdout.strip().splitlines()[-1])
    assert s["stamps"] >= 30
    for mesh in (s["heart"], s["after"]):
        assert mesh["watertight"] and mesh["loops"] >= 1 and mesh["overflow"] == 0
        assert mesh["surface_to_mesh_mean"] < 3.0
    assert s["after"]["loops"] >= 2  # the punched hole is its own boundary loop
