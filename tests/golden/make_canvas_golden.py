"""Writes tests/golden/canvas_heart.npz from the reference's own canvas.

The reference's canvas (src/voromesh/canvas.cpp) and heart outline
(src/voromesh/shape2d.cpp) are plain C++20 and compile in place:
`make -C oracle _ref/libref_canvas.so` builds them with a C shim
(oracle/ref_canvas_shim.cpp).  This script drives that library through the
demo's heart (apps/voromesh_demo.cpp:104: MakeHeart(512, 512, 819.2, 4096) in a
1024 x 1024 domain) at a 256-voxel raster, then the demo's scripted drags
(apps/voromesh_demo.cpp:535-541: stamps every quarter radius) and 40 seeded
random discs, and records the canvas after rasterization, every edit (fields
and dirty slots, in order) and the final target slots and occupancy masks.

Run:  make -C oracle _ref/libref_canvas.so && python tests/golden/make_canvas_golden.py
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_canvas.so")

DOMAIN, RASTER = 1024.0, 256
DRAGS = [((640, 120), (830, 230), 48, False), ((512, 560), (512, 560), 70, False), ((120, 300), (80, 440), 40, True)]


def ref_lib(path=REF_LIB):
    R = C.CDLL(path)
    vp = C.c_void_p
    R.ref_canvas_create.restype = vp
    R.ref_canvas_create.argtypes = [C.c_float] * 3
    R.ref_canvas_destroy.argtypes = [vp]
    R.ref_canvas_rasterize.argtypes = [vp, vp, C.c_int]
    R.ref_canvas_stamp.argtypes = [vp, C.c_float, C.c_float, C.c_float, C.c_int, vp, vp, C.c_int]
    R.ref_canvas_slots.argtypes = [vp, vp, C.c_int64]
    R.ref_canvas_slots.restype = C.c_int64
    R.ref_canvas_masks.argtypes = [vp, vp, C.c_int64]
    R.ref_canvas_masks.restype = C.c_int64
    R.ref_canvas_counts.argtypes = [vp, vp]
    R.ref_make_heart.argtypes = [C.c_float] * 3 + [C.c_int, vp]
    return R


def heart(R, cx, cy, width, segments):
    out = np.zeros((segments, 2), np.float32)
    R.ref_make_heart(cx, cy, width, segments, out.ctypes.data)
    return out


def ref_state(R, c):
    cnt = np.zeros(5, np.int64)
    R.ref_canvas_counts(c, cnt.ctypes.data)
    slots = np.zeros((cnt[0], 2), np.float32)
    R.ref_canvas_slots(c, slots.ctypes.data, cnt[0])
    masks = np.zeros(cnt[3] * cnt[4], np.uint64)
    R.ref_canvas_masks(c, masks.ctypes.data, len(masks))
    return cnt, slots, masks


def ref_stamp(R, c, x, y, r, add):
    edit = np.zeros(7, np.int32)
    dirty = np.zeros(1 << 20, np.int32)
    n = R.ref_canvas_stamp(c, x, y, r, int(add), edit.ctypes.data, dirty.ctypes.data, len(dirty))
    return edit, dirty[:n].copy()


def stamp_script(seed=5, n_random=40):
    """The demo's drags (centres every quarter radius, float32 like the demo) and random discs."""
    out = []
    for (a, b, r, add) in DRAGS:
        length = float(np.hypot(np.float32(b[0] - a[0]), np.float32(b[1] - a[1])))
        n = max(1, int(np.ceil(length / (0.25 * r))))
        for k in range(n + 1):
            t = np.float32(k) / np.float32(n)
            x = np.float32(a[0]) + t * np.float32(b[0] - a[0])
            y = np.float32(a[1]) + t * np.float32(b[1] - a[1])
            out.append((float(x), float(y), float(r), bool(add)))
            if length == 0.0:
                break
    rng = np.random.default_rng(seed)
    for _ in range(n_random):
        out.append((float(np.float32(rng.uniform(-40, DOMAIN + 40))), float(np.float32(rng.uniform(-40, DOMAIN + 40))),
                    float(np.float32(rng.uniform(3, 90))), bool(rng.integers(2))))
    return out


def main():
    R = ref_lib()
    voxel = DOMAIN / RASTER
    poly = heart(R, 0.5 * DOMAIN, 0.5 * DOMAIN, 0.8 * DOMAIN, 4096)
    c = R.ref_canvas_create(DOMAIN, DOMAIN, voxel)
    R.ref_canvas_rasterize(c, poly.ctypes.data, len(poly))
    cnt0, slots0, masks0 = ref_state(R, c)
    stamps = stamp_script()
    edits, dirty, offs = [], [], [0]
    for (x, y, r, add) in stamps:
        e, d = ref_stamp(R, c, x, y, r, add)
        edits.append(e)
        dirty.append(d)
        offs.append(offs[-1] + len(d))
    cnt1, slots1, masks1 = ref_state(R, c)
    R.ref_canvas_destroy(c)
    np.savez_compressed(
        os.path.join(HERE, "canvas_heart.npz"), domain=np.float32(DOMAIN), voxel=np.float32(voxel), polygon=poly,
        counts0=cnt0, slots0=slots0, masks0=masks0, stamps=np.array([s[:3] for s in stamps], np.float32),
        stamp_add=np.array([s[3] for s in stamps], np.uint8), edits=np.array(edits, np.int32),
        dirty=np.concatenate(dirty).astype(np.int32), dirty_offsets=np.array(offs, np.int64), counts1=cnt1,
        slots1=slots1, masks1=masks1)
    print(f"canvas_heart.npz: {cnt0[0]} slots after rasterization, {len(stamps)} stamps, {cnt1[0]} slots after")


if __name__ == "__main__":
    main()
