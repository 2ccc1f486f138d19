"""Python front end of libvoromesh_hip.so (ctypes over the C ABI in include/voromesh_hip.h).

Mirrors the reference's C++ interfaces for the VoroMesh path:
  * ``Optimizer``  -- voromesh::VoroOptimizer (include/voromesh/voro_optimizer.h:70-138)
  * ``Sampler``    -- poisson::Sampler (include/poisson/poisson_disk.h:21-62)
and adds ``Batch`` (many fits per launch) and ``poisson_batch``.

The library is required: importing this module raises if it has not been built
(``python -c "import __graft_entry__ as g; g.build()"``).  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "libvoromesh_hip.so"))

MAX_CELL_VERTS = 24


class VmError(RuntimeError):
    pass


class Vec2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class Params(C.Structure):
    """vm_params: voromesh::OptimizerParams plus the reference's compile-time knobs."""
    _fields_ = [("iterations", C.c_int), ("local_iterations", C.c_int), ("lr_per_spacing", C.c_float),
                ("sample_fraction", C.c_float), ("gamma", C.c_float), ("milestones", C.c_int * 4),
                ("seed", C.c_uint), ("loss_k", C.c_int), ("grid_cell", C.c_float), ("stage_rings", C.c_int),
                ("precision", C.c_int)]


class Mesh(C.Structure):
    _fields_ = [("num_slots", C.c_int), ("generators", C.POINTER(Vec2)), ("inside", C.POINTER(C.c_uint8)),
                ("cell_start", C.POINTER(C.c_int32)), ("cell_verts", C.POINTER(Vec2)),
                ("cell_labels", C.POINTER(C.c_int32)), ("num_vertices", C.c_int),
                ("vertices", C.POINTER(Vec2)), ("num_edges", C.c_int), ("edges", C.POINTER(C.c_int32)),
                ("watertight", C.c_int), ("loops", C.c_int), ("unsafe_cells", C.c_int),
                ("overflow_cells", C.c_int), ("gpu_ms", C.c_float), ("host_ms", C.c_float)]


class FitDesc(C.Structure):
    _fields_ = [("width", C.c_float), ("height", C.c_float), ("sites", C.c_void_p), ("num_sites", C.c_int),
                ("targets", C.c_void_p), ("num_targets", C.c_int), ("lr", C.c_float), ("gamma", C.c_float),
                ("milestones", C.c_int * 4), ("sample_fraction", C.c_float), ("seed", C.c_uint),
                ("loss_k", C.c_int), ("grid_cell", C.c_float), ("stage_rings", C.c_int),
                ("max_iterations", C.c_int), ("tol", C.c_float), ("chunk_extent", C.c_float),
                ("targets_per_lane", C.c_int)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_resident", C.c_int), ("max_iterations", C.c_int),
                ("tol", C.c_float), ("poll_every", C.c_int), ("precision", C.c_int),
                ("max_resident_sites", C.c_double), ("initial_pool_bytes", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [("final_sites", C.c_void_p), ("latency_s", C.c_double), ("queue_s", C.c_double), ("iterations", C.c_int),
                ("final_mean_loss", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("site_iters", C.c_double), ("site_iters_run", C.c_double), ("completed", C.c_int),
                ("max_resident_seen", C.c_int), ("peak_bytes", C.c_double), ("arrived", C.c_int),
                ("admitted", C.c_int), ("pool_grows", C.c_int), ("pool_bytes", C.c_double),
                ("gpu_busy_s", C.c_double), ("oldest_wait_s", C.c_double), ("prep_busy_s", C.c_double),
                ("sample_busy_s", C.c_double), ("loop_busy_s", C.c_double), ("prep_wait_s", C.c_double),
                ("upload_bytes", C.c_double),
                ("prep_workers", C.c_int)]


class ServeStream(C.Structure):
    _fields_ = [("rate", C.c_double), ("duration_s", C.c_double), ("seed", C.c_uint64), ("min_sites", C.c_int),
                ("max_sites", C.c_int), ("targets_per_site", C.c_int), ("sigma", C.c_float), ("lr", C.c_float),
                ("loss_k", C.c_int), ("drain", C.c_int)]


class ServeRecord(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double), ("num_sites", C.c_int),
                ("num_targets", C.c_int), ("iterations", C.c_int), ("final_mean_loss", C.c_double)]


# Every symbol include/voromesh_hip.h declares (checked by the CPU test suite).
EXPORTS = [
    "vm_default_params", "vm_last_error", "vm_version", "vm_device_count", "vm_copy_probe",
    "vm_create", "vm_destroy", "vm_set_targets", "vm_update_targets", "vm_set_occupancy", "vm_update_occupancy",
    "vm_set_generators", "vm_update_generators", "vm_begin_schedule", "vm_step", "vm_get_state",
    "vm_evaluate_loss", "vm_extract", "vm_get_generators", "vm_knn", "vm_loss_pass", "vm_graph_stats",
    "vm_grad_saturation", "vm_batch_saturation",
    "vm_sampler_create", "vm_sampler_destroy", "vm_sampler_resample", "vm_sampler_verify", "vm_sampler_dims",
    "vm_sampler_cells", "vm_sampler_cells_region", "vm_sampler_load_cells", "vm_poisson_grid_dims", "vm_poisson_batch",
    "vm_poisson_batch_noalloc",
    "vm_batch_create", "vm_batch_destroy", "vm_batch_add", "vm_batch_commit", "vm_batch_step",
    "vm_batch_get_sites", "vm_batch_status", "vm_batch_extract", "vm_batch_choice", "vm_batch_profile", "vm_batch_loss_stats",
    "vm_batch_loss_pass",
    "vm_serve", "vm_serve_synthetic", "vm_serve_synth_plan", "vm_serve_synth_targets",
    "vm_canvas_create", "vm_canvas_destroy", "vm_canvas_rasterize", "vm_canvas_stamp_disc", "vm_canvas_dirty",
    "vm_canvas_dims", "vm_canvas_counts", "vm_canvas_slots", "vm_canvas_masks", "vm_canvas_voxel",
]

_lib = None


class CanvasEdit(C.Structure):
    """vm_canvas_edit: occupancy tiles rewritten, samples removed/added, voxels flipped, dirty slots."""

    _fields_ = [(n, C.c_int32) for n in ("tx0", "ty0", "tx1", "ty1", "removed", "added", "flipped", "num_dirty")]


def lib():
    """Loads the HIP library (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VmError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    vp, i32, f32, f64 = C.c_void_p, C.c_int, C.c_float, C.c_double
    P = C.POINTER
    L.vm_last_error.restype = C.c_char_p
    L.vm_version.restype = C.c_char_p
    L.vm_default_params.argtypes = [P(Params)]
    L.vm_device_count.restype = i32
    sig = {
        "vm_copy_probe": [C.c_int64, P(f32)],
        "vm_create": [f32, f32, f32, P(Params), P(vp)],
        "vm_destroy": [vp],
        "vm_set_targets": [vp, vp, i32],
        "vm_update_targets": [vp, vp, i32, vp, i32],
        "vm_set_occupancy": [vp, vp, i32, i32, f32],
        "vm_update_occupancy": [vp, vp, i32, i32, i32, i32],
        "vm_set_generators": [vp, vp, i32, i32, f32],
        "vm_update_generators": [vp, vp, i32, i32, i32, i32],
        "vm_begin_schedule": [vp, i32, i32, i32, i32, i32],
        "vm_step": [vp, i32, P(f32)],
        "vm_get_state": [vp, P(i32), P(i32), P(f32), P(i32)],
        "vm_evaluate_loss": [vp, P(f64)],
        "vm_extract": [vp, P(Mesh)],
        "vm_get_generators": [vp, vp, i32],
        "vm_knn": [vp, vp, i32, i32, vp, vp],
        "vm_loss_pass": [vp, vp, vp, vp, vp, vp, vp, vp],
        "vm_graph_stats": [vp, P(C.c_int64), P(C.c_int64)],
        "vm_sampler_create": [f32, f32, f32, i32, C.c_uint, P(vp)],
        "vm_sampler_destroy": [vp],
        "vm_sampler_resample": [vp, i32, i32, i32, i32, P(f32)],
        "vm_sampler_verify": [vp, i32, i32, i32, i32, P(i32)],
        "vm_sampler_dims": [vp, P(i32), P(i32), P(f32)],
        "vm_sampler_cells": [vp, vp, i32],
        "vm_sampler_cells_region": [vp, i32, i32, i32, i32, vp, i32],
        "vm_sampler_load_cells": [vp, vp, i32],
        "vm_poisson_grid_dims": [f32, f32, f32, P(i32), P(i32), P(f32)],
        "vm_poisson_batch": [i32, vp, vp, vp, vp, i32, vp, vp, P(f32)],
        "vm_poisson_batch_noalloc": [i32, vp, vp, vp, vp, i32, vp, vp, P(f32)],
        "vm_batch_create": [i32, P(vp)],
        "vm_batch_destroy": [vp],
        "vm_batch_add": [vp, P(FitDesc), P(i32)],
        "vm_batch_commit": [vp],
        "vm_batch_step": [vp, i32, P(f32)],
        "vm_batch_get_sites": [vp, i32, vp, i32],
        "vm_batch_status": [vp, i32, P(i32), P(i32), P(f64)],
        "vm_batch_extract": [vp, i32, vp, i32, i32, f32, P(Mesh)],
        "vm_batch_choice": [vp, i32, P(i32), P(f32), P(i32), P(f32), P(i32)],
        "vm_batch_profile": [vp, i32, vp],
        "vm_batch_saturation": [vp, P(C.c_int64)],
        "vm_grad_saturation": [vp, P(C.c_int64)],
        "vm_batch_loss_stats": [vp, vp],
        "vm_batch_loss_pass": [vp, vp, vp, vp, vp, vp, vp],
        "vm_serve": [P(FitDesc), vp, i32, P(ServeConfig), vp, P(ServeStats)],
        "vm_serve_synthetic": [P(ServeStream), P(ServeConfig), vp, i32, P(i32), P(ServeStats)],
        "vm_serve_synth_plan": [P(ServeStream), C.c_int64, P(f64), P(i32), P(f32), P(C.c_uint)],
        "vm_serve_synth_targets": [P(ServeStream), C.c_int64, i32, vp, i32],
        "vm_canvas_create": [f32, f32, f32, P(vp)],
        "vm_canvas_destroy": [vp],
        "vm_canvas_rasterize": [vp, vp, i32],
        "vm_canvas_stamp_disc": [vp, Vec2, f32, i32, P(CanvasEdit)],
        "vm_canvas_dirty": [vp, vp, i32],
        "vm_canvas_dims": [vp, P(i32), P(i32), P(i32), P(i32), P(f32)],
        "vm_canvas_counts": [vp, P(C.c_int64), P(C.c_int64), P(C.c_int64)],
        "vm_canvas_slots": [vp, vp, C.c_int64],
        "vm_canvas_masks": [vp, vp, C.c_int64],
        "vm_canvas_voxel": [vp, i32, i32, P(i32), P(i32)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise VmError(f"status {rc}: {lib().vm_last_error().decode(errors='replace')}")


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _mesh_dict(m):
    """Copies a vm_mesh into numpy arrays."""
    n = m.num_slots

    def arr(p, count, dt, shape=None):
        if count == 0 or not p:
            return np.zeros(shape if shape else (0,), dt)
        a = np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(count,))
        a = a.copy()
        return a.reshape(shape) if shape else a

    start = arr(m.cell_start, n + 1, np.int32)
    nv = int(start[-1]) if n else 0
    return dict(
        generators=arr(C.cast(m.generators, C.POINTER(C.c_float)), 2 * n, np.float32, (n, 2)),
        inside=arr(m.inside, n, np.uint8), cell_start=start,
        cell_verts=arr(C.cast(m.cell_verts, C.POINTER(C.c_float)), 2 * nv, np.float32, (nv, 2)),
        cell_labels=arr(m.cell_labels, nv, np.int32),
        vertices=arr(C.cast(m.vertices, C.POINTER(C.c_float)), 2 * m.num_vertices, np.float32,
                     (m.num_vertices, 2)),
        edges=arr(m.edges, 2 * m.num_edges, np.int32, (m.num_edges, 2)),
        watertight=bool(m.watertight), loops=m.loops, unsafe_cells=m.unsafe_cells,
        overflow_cells=m.overflow_cells, gpu_ms=m.gpu_ms, host_ms=m.host_ms)


def _pts(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 2))


def device_count() -> int:
    return int(lib().vm_device_count())


def copy_probe(nbytes: int) -> float:
    """Device-to-device copy of nbytes (traffic-counter calibration); returns GPU ms."""
    ms = C.c_float()
    _check(lib().vm_copy_probe(int(nbytes), C.byref(ms)))
    return ms.value


def default_params(**kw) -> Params:
    p = Params()
    lib().vm_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "milestones":
            for i in range(4):
                p.milestones[i] = int(v[i])
        else:
            setattr(p, k, v)
    return p


class Optimizer:
    """voromesh::VoroOptimizer over the HIP library (same methods, same semantics)."""

    def __init__(self, width, height, spacing, params: Params | None = None, **kw):
        self._h = C.c_void_p()
        self.params = params if params is not None else default_params(**kw)
        _check(lib().vm_create(width, height, spacing, C.byref(self.params), C.byref(self._h)))
        self.width, self.height, self.spacing = width, height, spacing
        self.slots_w = self.slots_h = 0

    def close(self):
        if self._h:
            lib().vm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_targets(self, slots):
        s = _pts(slots)
        _check(lib().vm_set_targets(self._h, _ptr(s), len(s)))

    def update_targets(self, dirty, slots):
        d = np.ascontiguousarray(dirty, np.int32)
        s = _pts(slots)
        _check(lib().vm_update_targets(self._h, _ptr(d), len(d), _ptr(s), len(s)))

    def set_occupancy(self, masks, tiles_x, tiles_y, voxel):
        m = np.ascontiguousarray(masks, np.uint64)
        _check(lib().vm_set_occupancy(self._h, _ptr(m), tiles_x, tiles_y, voxel))

    def update_occupancy(self, masks, tx0, ty0, tx1, ty1):
        m = np.ascontiguousarray(masks, np.uint64)
        _check(lib().vm_update_occupancy(self._h, _ptr(m), tx0, ty0, tx1, ty1))

    def set_generators(self, slots, slots_w, slots_h, slot_cell):
        s = _pts(slots)
        assert len(s) == slots_w * slots_h
        self.slots_w, self.slots_h = slots_w, slots_h
        _check(lib().vm_set_generators(self._h, _ptr(s), slots_w, slots_h, slot_cell))

    def update_generators(self, slots, cx0, cy0, cx1, cy1):
        s = _pts(slots)
        _check(lib().vm_update_generators(self._h, _ptr(s), cx0, cy0, cx1, cy1))

    def begin_schedule(self, iterations, rect=None):
        if rect is None:
            rect = (0, 0, self.slots_w, self.slots_h)
        _check(lib().vm_begin_schedule(self._h, iterations, *rect))

    def step(self, iterations) -> float:
        ms = C.c_float()
        _check(lib().vm_step(self._h, iterations, C.byref(ms)))
        return ms.value

    def state(self):
        it, sl, ls, ns = C.c_int(), C.c_int(), C.c_float(), C.c_int()
        _check(lib().vm_get_state(self._h, C.byref(it), C.byref(sl), C.byref(ls), C.byref(ns)))
        return dict(iteration=it.value, schedule_length=sl.value, last_batch_loss=ls.value, num_slots=ns.value)

    def optimize(self) -> float:
        st = self.state()
        return self.step(st["schedule_length"] - st["iteration"])

    def graph_stats(self):
        """Iteration-graph captures taken in place by an existing exec, and execs instantiated."""
        u, i = C.c_int64(), C.c_int64()
        _check(lib().vm_graph_stats(self._h, C.byref(u), C.byref(i)))
        return dict(updates=u.value, instantiations=i.value)

    def grad_saturation(self):
        """Targets whose FP32 fixed-point gradient terms were clamped so far (vm_grad_saturation)."""
        n = C.c_int64()
        _check(lib().vm_grad_saturation(self._h, C.byref(n)))
        return n.value

    def evaluate_loss(self) -> float:
        v = C.c_double()
        _check(lib().vm_evaluate_loss(self._h, C.byref(v)))
        return v.value

    def generators(self):
        n = self.slots_w * self.slots_h
        out = np.zeros((n, 2), np.float32)
        _check(lib().vm_get_generators(self._h, _ptr(out), n))
        return out

    def knn(self, queries, k):
        q = _pts(queries)
        idx = np.zeros((len(q), k), np.int32)
        d2 = np.zeros((len(q), k), np.float32)
        _check(lib().vm_knn(self._h, _ptr(q), len(q), k, _ptr(idx), _ptr(d2)))
        return idx, d2

    def loss_pass(self, n_target_slots, want_grad=True, want_terms=True):
        n = self.slots_w * self.slots_h
        K = self.params.loss_k
        grad = np.zeros((n, 2), np.float64) if want_grad else None
        ls, cnt = np.zeros(1, np.float64), np.zeros(1, np.int64)
        knn = near = part = terms = None
        if want_terms:
            knn = np.zeros((n_target_slots, K), np.int32)
            near = np.zeros(n_target_slots, np.int32)
            part = np.zeros(n_target_slots, np.int32)
            terms = np.zeros(n_target_slots, np.float64)
        _check(lib().vm_loss_pass(self._h, _ptr(grad), _ptr(ls), _ptr(cnt), _ptr(knn), _ptr(near), _ptr(part),
                                  _ptr(terms)))
        return dict(grad=grad, loss_sum=float(ls[0]), count=int(cnt[0]), knn=knn, nearest=near, partner=part,
                    loss_terms=terms)

    def extract(self):
        m = Mesh()
        _check(lib().vm_extract(self._h, C.byref(m)))
        return _mesh_dict(m)


class Canvas:
    """voromesh::Canvas (host voxel canvas): target slots and occupancy masks from a
    polygon and disc edits (include/voromesh/canvas.h:27-89 of the reference)."""

    def __init__(self, width, height, voxel_size):
        self._c = C.c_void_p()
        _check(lib().vm_canvas_create(width, height, voxel_size, C.byref(self._c)))
        vx, vy, tx, ty, v = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_float()
        _check(lib().vm_canvas_dims(self._c, C.byref(vx), C.byref(vy), C.byref(tx), C.byref(ty), C.byref(v)))
        self.voxels_x, self.voxels_y, self.tiles_x, self.tiles_y = vx.value, vy.value, tx.value, ty.value
        self.voxel_size = v.value

    def close(self):
        if self._c:
            lib().vm_canvas_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rasterize(self, polygon):
        poly = _pts(polygon)
        _check(lib().vm_canvas_rasterize(self._c, _ptr(poly), len(poly)))

    def stamp_disc(self, center, radius, inside):
        """Returns (edit dict, dirty slots int32 array)."""
        e = CanvasEdit()
        _check(lib().vm_canvas_stamp_disc(self._c, Vec2(float(center[0]), float(center[1])), float(radius),
                                          1 if inside else 0, C.byref(e)))
        dirty = np.zeros(max(e.num_dirty, 0), np.int32)
        _check(lib().vm_canvas_dirty(self._c, _ptr(dirty), len(dirty)))
        return {n: getattr(e, n) for n, _ in CanvasEdit._fields_}, dirty

    def counts(self):
        n, s, t = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().vm_canvas_counts(self._c, C.byref(n), C.byref(s), C.byref(t)))
        return n.value, s.value, t.value

    def target_slots(self):
        out = np.zeros((self.counts()[0], 2), np.float32)
        _check(lib().vm_canvas_slots(self._c, _ptr(out), len(out)))
        return out

    def inside_masks(self):
        out = np.zeros(self.tiles_x * self.tiles_y, np.uint64)
        _check(lib().vm_canvas_masks(self._c, _ptr(out), len(out)))
        return out

    def voxel(self, i, j):
        a, b = C.c_int(), C.c_int()
        _check(lib().vm_canvas_voxel(self._c, i, j, C.byref(a), C.byref(b)))
        return bool(a.value), bool(b.value)


class Sampler:
    """poisson::Sampler over the HIP dart thrower."""

    def __init__(self, width, height, radius, trials=30, seed=1):
        self._s = C.c_void_p()
        _check(lib().vm_sampler_create(width, height, radius, trials, seed, C.byref(self._s)))
        gw, gh, cs = C.c_int(), C.c_int(), C.c_float()
        _check(lib().vm_sampler_dims(self._s, C.byref(gw), C.byref(gh), C.byref(cs)))
        self.grid_w, self.grid_h, self.cell_size = gw.value, gh.value, cs.value
        self.width, self.height, self.radius = width, height, radius

    def close(self):
        if self._s:
            lib().vm_sampler_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def resample(self, cx0, cy0, cx1, cy1) -> float:
        ms = C.c_float()
        _check(lib().vm_sampler_resample(self._s, cx0, cy0, cx1, cy1, C.byref(ms)))
        return ms.value

    def verify(self, cx0=0, cy0=0, cx1=None, cy1=None) -> bool:
        ok = C.c_int()
        cx1 = self.grid_w if cx1 is None else cx1
        cy1 = self.grid_h if cy1 is None else cy1
        _check(lib().vm_sampler_verify(self._s, cx0, cy0, cx1, cy1, C.byref(ok)))
        return bool(ok.value)

    def cells(self):
        n = self.grid_w * self.grid_h
        out = np.zeros((n, 2), np.float32)
        _check(lib().vm_sampler_cells(self._s, _ptr(out), n))
        return out

    def load_cells(self, cells):
        """Replaces the grid (host mirror and device copy); later resamples continue from it."""
        c = _pts(cells)
        _check(lib().vm_sampler_load_cells(self._s, _ptr(c), len(c)))


def poisson_grid_dims(width, height, radius):
    gw, gh, cs = C.c_int(), C.c_int(), C.c_float()
    _check(lib().vm_poisson_grid_dims(width, height, radius, C.byref(gw), C.byref(gh), C.byref(cs)))
    return gw.value, gh.value, cs.value


def poisson_batch(widths, heights, radii, seeds, trials=30, noalloc=False):
    """Samples many Poisson-disk grids in one set of launches; returns (list of (gh*gw,2) grids, dims, gpu_ms).
    noalloc: through vm_poisson_batch_noalloc (fixed device array, no allocation; bounded sizes)."""
    n = len(radii)
    w = np.ascontiguousarray(widths, np.float32)
    h = np.ascontiguousarray(heights, np.float32)
    r = np.ascontiguousarray(radii, np.float32)
    s = np.ascontiguousarray(seeds, np.uint32)
    dims = [poisson_grid_dims(float(w[i]), float(h[i]), float(r[i])) for i in range(n)]
    offs = np.zeros(n, np.int64)
    total = 0
    for i, (gw, gh, _) in enumerate(dims):
        offs[i] = total
        total += gw * gh
    cells = np.zeros((max(total, 1), 2), np.float32)
    ms = C.c_float()
    fn = lib().vm_poisson_batch_noalloc if noalloc else lib().vm_poisson_batch
    _check(fn(n, _ptr(w), _ptr(h), _ptr(r), _ptr(s), trials, _ptr(offs), _ptr(cells), C.byref(ms)))
    grids = [cells[offs[i]:offs[i] + dims[i][0] * dims[i][1]] for i in range(n)]
    return grids, dims, ms.value


def fit_desc(width, height, sites, targets, lr, milestones=(80, 150, 200, 250), gamma=0.5, sample_fraction=1.0,
             seed=1, loss_k=16, grid_cell=0.0, stage_rings=0, max_iterations=0, tol=0.0, chunk_extent=0.0,
             targets_per_lane=0):
    """A FitDesc plus the arrays it points to (keep the pair alive together)."""
    s, t = _pts(sites), _pts(targets)
    d = FitDesc()
    d.width, d.height = width, height
    d.sites, d.num_sites = s.ctypes.data, len(s)
    d.targets, d.num_targets = t.ctypes.data, len(t)
    d.lr, d.gamma = lr, gamma
    for i in range(4):
        d.milestones[i] = int(milestones[i])
    d.sample_fraction, d.seed, d.loss_k = sample_fraction, seed, loss_k
    d.grid_cell, d.stage_rings, d.max_iterations, d.tol = grid_cell, stage_rings, max_iterations, tol
    d.chunk_extent = chunk_extent
    d.targets_per_lane = targets_per_lane
    return d, (s, t)


class Batch:
    """Many independent fits advanced by one set of launches per iteration."""

    def __init__(self, precision=0):
        self._b = C.c_void_p()
        _check(lib().vm_batch_create(precision, C.byref(self._b)))
        self._keep = []
        self.sizes = []
        self.target_counts = []

    def close(self):
        if self._b:
            lib().vm_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, width, height, sites, targets, lr, **kw) -> int:
        d, keep = fit_desc(width, height, sites, targets, lr, **kw)
        i = C.c_int()
        _check(lib().vm_batch_add(self._b, C.byref(d), C.byref(i)))
        self._keep.append(keep)
        self.sizes.append(int((keep[0][:, 0] >= 0).sum()))
        self.target_counts.append(int((keep[1][:, 0] >= 0).sum()))
        return i.value

    def commit(self):
        _check(lib().vm_batch_commit(self._b))
        self._keep = []

    def step(self, iterations) -> float:
        ms = C.c_float()
        _check(lib().vm_batch_step(self._b, iterations, C.byref(ms)))
        return ms.value

    def profile(self, iterations):
        """Mean ms per phase: scan, scatter, loss, exact selection, adam."""
        out = np.zeros(5, np.float32)
        _check(lib().vm_batch_profile(self._b, iterations, _ptr(out)))
        return out

    def saturation(self):
        """Clamped FP32 gradient terms summed over the batch's fits (vm_batch_saturation)."""
        n = C.c_int64()
        _check(lib().vm_batch_saturation(self._b, C.byref(n)))
        return n.value

    def sites(self, i):
        out = np.zeros((self.sizes[i], 2), np.float32)
        _check(lib().vm_batch_get_sites(self._b, i, _ptr(out), self.sizes[i]))
        return out

    def status(self, i):
        it, done, loss = C.c_int(), C.c_int(), C.c_double()
        _check(lib().vm_batch_status(self._b, i, C.byref(it), C.byref(done), C.byref(loss)))
        return dict(iteration=it.value, done=bool(done.value), last_mean_loss=loss.value)

    def loss_pass(self, want_terms=True):
        """One loss pass over every fit at its current sites: gradient sums and per-target terms, concatenated in
        the order the fits were added (site indices instance-local)."""
        nS = sum(self.sizes)
        nT = sum(self.target_counts)
        grad = np.zeros((nS, 2), np.float64)
        ls, cnt = np.zeros(1, np.float64), np.zeros(1, np.int64)
        near = part = terms = None
        if want_terms:
            near, part = np.zeros(nT, np.int32), np.zeros(nT, np.int32)
            terms = np.zeros(nT, np.float64)
        _check(lib().vm_batch_loss_pass(self._b, _ptr(grad), _ptr(ls), _ptr(cnt), _ptr(near), _ptr(part),
                                        _ptr(terms)))
        return dict(grad=grad, loss_sum=float(ls[0]), count=int(cnt[0]), nearest=near, partner=part,
                    loss_terms=terms)

    def loss_stats(self):
        c = np.zeros(8, np.uint64)
        _check(lib().vm_batch_loss_stats(self._b, _ptr(c)))
        keys = ["fast_waves", "ring_staged_waves", "ring_global_waves", "idle_waves", "candidates", "staged_sites",
                "exact_fallback_waves", "exact_fallback_targets"]
        return {k: int(v) for k, v in zip(keys, c)}

    def extract(self, i, masks=None, tiles=(0, 0), voxel=0.0):
        """Clipped Voronoi cells of fit i (CSR over its live sites) and, with occupancy masks, the boundary."""
        m = Mesh()
        mk = None if masks is None else np.ascontiguousarray(masks, np.uint64)
        _check(lib().vm_batch_extract(self._b, i, _ptr(mk), int(tiles[0]), int(tiles[1]), float(voxel), C.byref(m)))
        return _mesh_dict(m)

    def choice(self, i):
        k, cell, rings, ext, tpl = C.c_int(), C.c_float(), C.c_int(), C.c_float(), C.c_int()
        _check(lib().vm_batch_choice(self._b, i, C.byref(k), C.byref(cell), C.byref(rings), C.byref(ext),
                                     C.byref(tpl)))
        return dict(loss_k=k.value, grid_cell=cell.value, stage_rings=rings.value, chunk_extent=ext.value,
                    targets_per_lane=tpl.value)


def _serve_config(hbm_budget_bytes, max_resident, max_iterations, tol, poll_every, precision, max_resident_sites,
                  initial_pool_bytes):
    cfg = ServeConfig()
    cfg.hbm_budget_bytes, cfg.max_resident, cfg.max_iterations = hbm_budget_bytes, max_resident, max_iterations
    cfg.tol, cfg.poll_every, cfg.precision = tol, poll_every, precision
    cfg.max_resident_sites, cfg.initial_pool_bytes = max_resident_sites, initial_pool_bytes
    return cfg


def _stats_dict(st):
    return {name: getattr(st, name) for name, _ in ServeStats._fields_}


def serve(requests, arrivals, hbm_budget_bytes=200e9, max_resident=0, max_iterations=300, tol=1e-6, poll_every=10,
          precision=0, want_sites=False, max_resident_sites=0, initial_pool_bytes=0):
    """Continuous-batching request loop (vm_serve).

    requests: list of dicts with keys width, height, sites, targets, lr and optional fit_desc keywords.
    Returns (per-request dicts, stats dict).
    """
    n = len(requests)
    descs = (FitDesc * max(n, 1))()
    keep = []
    for i, r in enumerate(requests):
        kw = {k: v for k, v in r.items() if k not in ("width", "height", "sites", "targets", "lr")}
        d, k = fit_desc(r["width"], r["height"], r["sites"], r["targets"], r["lr"], **kw)
        descs[i] = d
        keep.append(k)
    arr = np.ascontiguousarray(arrivals, np.float64)
    assert len(arr) == n
    results = (ServeResult * max(n, 1))()
    bufs = []
    if want_sites:
        for i in range(n):
            b = np.zeros((descs[i].num_sites, 2), np.float32)
            bufs.append(b)
            results[i].final_sites = b.ctypes.data
    cfg = _serve_config(hbm_budget_bytes, max_resident, max_iterations, tol, poll_every, precision,
                        max_resident_sites, initial_pool_bytes)
    stats = ServeStats()
    _check(lib().vm_serve(descs, _ptr(arr), n, C.byref(cfg), C.cast(results, C.c_void_p), C.byref(stats)))
    out = []
    for i in range(n):
        r = results[i]
        o = dict(latency_s=r.latency_s, queue_s=r.queue_s, iterations=r.iterations,
                 final_mean_loss=r.final_mean_loss)
        if want_sites:
            o["sites"] = bufs[i]
        out.append(o)
    return out, _stats_dict(stats)


def serve_stream_desc(rate, duration_s, seed=4, min_sites=512, max_sites=131072, targets_per_site=8, sigma=5e-4,
                      lr=1e-3, loss_k=16, drain=False):
    """A vm_serve_stream: the synthetic request stream the library generates lazily."""
    st = ServeStream()
    st.rate, st.duration_s, st.seed = rate, duration_s, seed
    st.min_sites, st.max_sites, st.targets_per_site = min_sites, max_sites, targets_per_site
    st.sigma, st.lr, st.loss_k, st.drain = sigma, lr, loss_k, int(bool(drain))
    return st


RECORD_DTYPE = np.dtype([("arrival_s", np.float64), ("admit_s", np.float64), ("done_s", np.float64),
                         ("num_sites", np.int32), ("num_targets", np.int32), ("iterations", np.int32),
                         ("final_mean_loss", np.float64)], align=True)


def serve_synthetic(stream, hbm_budget_bytes=200e9, max_resident=0, max_iterations=300, tol=1e-6, poll_every=10,
                    precision=0, max_resident_sites=0, initial_pool_bytes=0, max_records=None):
    """Serves a synthetic stream (vm_serve_synthetic); returns (records array of RECORD_DTYPE, stats dict)."""
    assert RECORD_DTYPE.itemsize == C.sizeof(ServeRecord)
    if max_records is None:
        mean = stream.rate * stream.duration_s
        max_records = int(mean + 10 * mean ** 0.5 + 64)
    rec = np.zeros(max_records, RECORD_DTYPE)
    cfg = _serve_config(hbm_budget_bytes, max_resident, max_iterations, tol, poll_every, precision,
                        max_resident_sites, initial_pool_bytes)
    n = C.c_int()
    stats = ServeStats()
    _check(lib().vm_serve_synthetic(C.byref(stream), C.byref(cfg), _ptr(rec), max_records, C.byref(n),
                                    C.byref(stats)))
    return rec[:n.value].copy(), _stats_dict(stats)


def synth_plan(stream, index):
    """(arrival_s, drawn site count, Poisson-disk radius, sampler seed) of request `index` of a stream."""
    t, s, r, sd = C.c_double(), C.c_int(), C.c_float(), C.c_uint()
    _check(lib().vm_serve_synth_plan(C.byref(stream), index, C.byref(t), C.byref(s), C.byref(r), C.byref(sd)))
    return t.value, s.value, r.value, sd.value


def synth_targets(stream, index, num_sites):
    """Request `index`'s targets for its num_sites live sites."""
    out = np.zeros((stream.targets_per_site * num_sites, 2), np.float32)
    _check(lib().vm_serve_synth_targets(C.byref(stream), index, num_sites, _ptr(out), len(out)))
    return out


def synth_request(stream, index):
    """Request `index` of a synthetic stream as the server generates it: (dict(width, height, sites, targets, lr,
    loss_k) for vm_serve, arrival time)."""
    t, _, r, sd = synth_plan(stream, index)
    grids, _, _ = poisson_batch([1.0], [1.0], [r], [sd])
    g = grids[0]
    sites = g[g[:, 0] >= 0]
    return dict(width=1.0, height=1.0, sites=sites, targets=synth_targets(stream, index, len(sites)), lr=stream.lr,
                loss_k=stream.loss_k or 16), t
