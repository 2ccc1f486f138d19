"""The voxel canvas (host code, no GPU): target slots, occupancy masks and disc
edits of include/voromesh/canvas.h / src/voromesh/canvas.cpp of the reference.

Parity is pinned three ways: against tests/golden/canvas_heart.npz, written by
the reference's own canvas (tests/golden/make_canvas_golden.py); against that
library live on random shapes and edits when oracle/_ref is built here; and
against what the geometry fixes (a brute-force even-odd test per voxel centre,
sample means inside their voxel, slot bookkeeping).  Everything is bit-exact.
"""
import ctypes as C
import os

import numpy as np
import pytest

import vmhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "canvas_heart.npz")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_canvas.so")
EDIT_FIELDS = ("tx0", "ty0", "tx1", "ty1", "removed", "added", "flipped")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_canvas_matches_reference_fixture():
    g = np.load(GOLDEN)
    c = vmhip.Canvas(float(g["domain"]), float(g["domain"]), float(g["voxel"]))
    c.rasterize(g["polygon"])
    assert c.counts() == tuple(int(v) for v in g["counts0"][:3])
    assert np.array_equal(bits(c.target_slots()), bits(g["slots0"]))
    assert np.array_equal(c.inside_masks(), g["masks0"])
    offs = g["dirty_offsets"]
    for k, (st, add) in enumerate(zip(g["stamps"], g["stamp_add"])):
        e, dirty = c.stamp_disc((st[0], st[1]), st[2], bool(add))
        assert [e[f] for f in EDIT_FIELDS] == list(g["edits"][k]), k
        assert e["num_dirty"] == offs[k + 1] - offs[k]
        assert np.array_equal(dirty, g["dirty"][offs[k]:offs[k + 1]]), k
    assert c.counts() == tuple(int(v) for v in g["counts1"][:3])
    assert np.array_equal(bits(c.target_slots()), bits(g["slots1"]))
    assert np.array_equal(c.inside_masks(), g["masks1"])


def _inside_brute(poly, W, H, voxel):
    """Even-odd test of every voxel centre against the polygon, as tile masks."""
    vx, vy = int(np.ceil(np.float32(W) / np.float32(voxel))), int(np.ceil(np.float32(H) / np.float32(voxel)))
    tx, ty = (vx + 7) // 8, (vy + 7) // 8
    masks = np.zeros(tx * ty, np.uint64)
    P = np.asarray(poly, np.float64)
    for j in range(vy):
        y = (j + 0.5) * voxel
        for i in range(vx):
            x = (i + 0.5) * voxel
            inside = False
            for s in range(len(P)):
                a, b = P[s], P[s - 1]
                if (a[1] > y) != (b[1] > y) and x < a[0] + (y - a[1]) * (b[0] - a[0]) / (b[1] - a[1]):
                    inside = not inside
            if inside:
                masks[(j >> 3) * tx + (i >> 3)] |= np.uint64(1) << np.uint64(((j & 7) << 3) | (i & 7))
    return masks


def test_rasterize_square_against_brute_force():
    # Vertices off the voxel-centre lattice, so float rounding cannot decide a centre.
    poly = np.array([[3.3, 2.7], [27.1, 4.2], [25.6, 29.4], [5.2, 26.9]], np.float32)
    c = vmhip.Canvas(32.0, 32.0, 1.0)
    c.rasterize(poly)
    assert np.array_equal(c.inside_masks(), _inside_brute(poly, 32.0, 32.0, 1.0))
    slots = c.target_slots()
    n, surf, tiles = c.counts()
    assert n == surf == len(slots) > 0
    # Every surface sample is the mean of outline samples inside its voxel.
    for (x, y) in slots:
        i, j = int(np.floor(x)), int(np.floor(y))
        assert c.voxel(i, j)[1], (x, y)


def test_edits_keep_slot_bookkeeping():
    c = vmhip.Canvas(64.0, 48.0, 1.0)
    c.rasterize(np.array([[10, 10], [50, 12], [40, 40], [12, 35]], np.float32))
    rng = np.random.default_rng(3)
    for _ in range(60):
        e, dirty = c.stamp_disc((rng.uniform(-5, 70), rng.uniform(-5, 55)), rng.uniform(0.5, 12), bool(rng.integers(2)))
        slots = c.target_slots()
        n, surf, _ = c.counts()
        live = int(np.count_nonzero(slots[:, 0] >= 0))
        assert live == surf and len(slots) == n
        assert np.all(dirty >= 0) and np.all(dirty < n)
        assert e["removed"] + e["added"] <= len(dirty) or len(dirty) == 0


def test_empty_and_out_of_domain():
    c = vmhip.Canvas(16.0, 16.0, 1.0)
    c.rasterize(np.zeros((0, 2), np.float32))
    assert c.counts() == (0, 0, 0) and not c.inside_masks().any()
    e, dirty = c.stamp_disc((-50.0, -50.0), 5.0, True)      # entirely outside: no-op
    assert all(e[f] == 0 for f in EDIT_FIELDS) and len(dirty) == 0
    e, dirty = c.stamp_disc((8.0, 8.0), 3.0, False)          # carve nothing out of nothing
    assert e["flipped"] == 0 and e["removed"] == 0 and e["added"] == 0
    e, dirty = c.stamp_disc((8.0, 8.0), 3.0, True)           # add a disc: rim becomes surface
    assert e["flipped"] > 0 and e["added"] > 0 and len(dirty) == e["added"]
    with pytest.raises(vmhip.VmError):
        vmhip.Canvas(0.0, 16.0, 1.0)


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference canvas (oracle/_ref) not built here")
def test_canvas_matches_reference_live():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mcg", os.path.join(ROOT, "tests", "golden", "make_canvas_golden.py"))
    mcg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mcg)
    R = mcg.ref_lib(REF_LIB)
    rng = np.random.default_rng(11)
    for case in range(6):
        W, H = float(rng.uniform(40, 300)), float(rng.uniform(40, 300))
        voxel = float(np.float32(rng.uniform(0.37, 3.1)))
        # a star polygon with random radii
        n = int(rng.integers(5, 400))
        t = np.sort(rng.uniform(0, 2 * np.pi, n))
        rad = rng.uniform(0.2, 0.5, n) * min(W, H)
        poly = np.stack([0.5 * W + rad * np.cos(t), 0.5 * H + rad * np.sin(t)], 1).astype(np.float32)
        rc = R.ref_canvas_create(W, H, voxel)
        R.ref_canvas_rasterize(rc, poly.ctypes.data, len(poly))
        c = vmhip.Canvas(W, H, voxel)
        c.rasterize(poly)
        for k in range(120):
            x, y = np.float32(rng.uniform(-0.1 * W, 1.1 * W)), np.float32(rng.uniform(-0.1 * H, 1.1 * H))
            r, add = np.float32(rng.uniform(0.2, 0.25 * min(W, H))), bool(rng.integers(2))
            edit, dirty = mcg.ref_stamp(R, rc, x, y, r, add)
            e, d = c.stamp_disc((x, y), r, add)
            assert [e[f] for f in EDIT_FIELDS] == list(edit), (case, k)
            assert np.array_equal(d, dirty), (case, k)
        cnt, slots, masks = mcg.ref_state(R, rc)
        assert c.counts() == tuple(int(v) for v in cnt[:3])
        assert np.array_equal(bits(c.target_slots()), bits(slots))
        assert np.array_equal(c.inside_masks(), masks)
        R.ref_canvas_destroy(rc)


HEART_DUMP = os.path.join(ROOT, "tests", "cpp", "heart_dump")


@pytest.mark.skipif(not os.path.exists(HEART_DUMP), reason="tests/cpp not built (run __graft_entry__.build())")
def test_make_heart_matches_reference_outline():
    import subprocess
    g = np.load(GOLDEN)
    out = subprocess.run([HEART_DUMP, "512", "512", "819.2", "4096"], capture_output=True, text=True, check=True)
    xy = np.array([[int(a, 16), int(b, 16)] for a, b in (ln.split() for ln in out.stdout.split("\n") if ln)], np.uint32)
    assert np.array_equal(xy, bits(g["polygon"]))
