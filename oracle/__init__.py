"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the tests, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` are the only users.  The product (the HIP library
and its host code) never imports this module.

Every function restates the reference's algorithm (src/voromesh/voro_optimizer.cu,
src/poisson/poisson_disk.cu); see voro_oracle.cpp for the line citations.
Point arrays are numpy arrays of shape (n, 2) in float32 (``real=0``) or
float64 (``real=1``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

KMAX_CELL_VERTS = 24


class Sched(C.Structure):
    """Adam + MultiStepLR schedule (voro_optimizer.cu:88-95 RunParams subset)."""
    _fields_ = [("lr", C.c_double), ("gamma", C.c_double), ("beta1", C.c_double),
                ("beta2", C.c_double), ("eps", C.c_double), ("milestones", C.c_int * 4),
                ("width", C.c_double), ("height", C.c_double)]


def build(force: bool = False) -> str:
    """Compiles liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp, i32, u32, f32, f64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_float, C.c_double
        L.vo_pcg_hash.restype = u32
        L.vo_pcg_hash.argtypes = [u32]
        L.vo_rand01_loss.restype = f32
        L.vo_rand01_loss.argtypes = [u32, u32, u32]
        L.vo_rand01_poisson.restype = f32
        L.vo_rand01_poisson.argtypes = [u32, u32, u32, u32]
        L.vo_num_threads.restype = C.c_int
        L.vo_set_threads.argtypes = [C.c_int]
        L.vo_knn.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, f64, f64, f64, C.c_int, vp, vp]
        L.vo_loss_grad.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int, f64, f64, f64, f32, u32, u32,
                                   vp, vp, vp, vp, vp, vp, vp, vp]
        L.vo_adam.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int64, C.POINTER(Sched)]
        L.vo_fit.argtypes = [C.c_int, vp, vp, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, f64, f32, u32, vp,
                             C.c_int, C.c_int, C.POINTER(Sched), vp, vp]
        L.vo_clip.argtypes = [C.c_int, vp, C.c_int, f64, f64, f64, C.c_int, vp, C.c_int, C.c_int, f32,
                              vp, vp, vp, vp, vp]
        L.vo_boundary.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_int, vp, C.c_int, vp, vp, vp, vp]
        L.vo_poisson_dims.argtypes = [f32, f32, f32, vp, vp, vp]
        L.vo_poisson.argtypes = [f32, f32, f32, C.c_int, u32, C.c_int, C.c_int, C.c_int, C.c_int, u32, vp]
        L.vo_poisson_verify.argtypes = [f32, f32, f32, vp, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _pts(a, real):
    dt = np.float64 if real else np.float32
    return np.ascontiguousarray(np.asarray(a, dtype=dt).reshape(-1, 2))


def set_threads(n: int) -> None:
    lib().vo_set_threads(int(n))


def num_threads() -> int:
    return int(lib().vo_num_threads())


def pcg_hash(v: int) -> int:
    return int(lib().vo_pcg_hash(v & 0xFFFFFFFF))


def rand01_loss(index: int, it: int, seed: int) -> float:
    return float(lib().vo_rand01_loss(index, it, seed))


def rand01_poisson(cell: int, rnd: int, seed: int, stream: int) -> float:
    return float(lib().vo_rand01_poisson(cell, rnd, seed, stream))


def knn(sites, queries, K, W, H, cell, real=0, brute=False):
    """k nearest sites of every query: (idx[Q,K] int32, d2[Q,K])."""
    s, q = _pts(sites, real), _pts(queries, real)
    idx = np.empty((len(q), K), np.int32)
    d2 = np.empty((len(q), K), np.float64 if real else np.float32)
    rc = lib().vo_knn(real, _p(s), len(s), _p(q), len(q), K, W, H, cell, int(brute), _p(idx), _p(d2))
    assert rc == 0
    return idx, d2


def loss_grad(sites, targets, K, W, H, cell, it=1, fraction=1.0, seed=1, box=None, slot_of=None, real=0,
              per_target=False):
    """One VoroLoss pass: dict(grad[S,2] float64 sums, loss_sum, count, and optional per-target terms)."""
    s, t = _pts(sites, real), _pts(targets, real)
    grad = np.zeros((len(s), 2), np.float64)
    loss_sum = np.zeros(1, np.float64)
    count = np.zeros(1, np.int64)
    boxa = None if box is None else np.asarray(box, np.float64)
    so = None if slot_of is None else np.ascontiguousarray(slot_of, np.int32)
    out = {}
    c0 = bj = lt = tg = None
    if per_target:
        c0 = np.empty(len(t), np.int32)
        bj = np.empty(len(t), np.int32)
        lt = np.empty(len(t), np.float64 if real else np.float32)
        tg = np.empty((len(t), 4), np.float64 if real else np.float32)
    rc = lib().vo_loss_grad(real, _p(s), len(s), _p(t), len(t), _p(so), K, W, H, cell, fraction, seed, it,
                            _p(boxa), _p(grad), _p(loss_sum), _p(count), _p(c0), _p(bj), _p(lt), _p(tg))
    assert rc == 0
    out.update(grad=grad, loss_sum=float(loss_sum[0]), count=int(count[0]))
    if per_target:
        out.update(c0=c0, bestj=bj, loss_t=lt, term_grad=tg)
    return out


def make_sched(lr, W, H, gamma=0.5, milestones=(80, 150, 200, 250), beta1=0.9, beta2=0.999, eps=1e-8):
    sc = Sched()
    sc.lr, sc.gamma, sc.beta1, sc.beta2, sc.eps = lr, gamma, beta1, beta2, eps
    for k in range(4):
        sc.milestones[k] = int(milestones[k])
    sc.width, sc.height = W, H
    return sc


def scale_milestones(milestones, iterations, schedule_iterations):
    """Milestones scaled to a schedule of `iterations` steps (voro_optimizer.cu:664-665)."""
    import math
    out = []
    for m in milestones:
        x = float(m) * iterations / schedule_iterations
        out.append(int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5)))
    return out


def adam(sites, m, v, grad, t, count, sched, movable=None, real=0):
    """In-place Adam step on float arrays (sites, m, v) with gradient sums `grad` (float64)."""
    g = np.ascontiguousarray(grad, np.float64)
    mv = None if movable is None else np.ascontiguousarray(movable, np.uint8)
    rc = lib().vo_adam(real, _p(sites), _p(m), _p(v), _p(g), _p(mv), len(sites), t, count, C.byref(sched))
    assert rc == 0


def fit(sites, targets, K, cell, sched, iters, t_start=1, fraction=1.0, seed=1, box=None, movable=None,
        slot_of=None, m=None, v=None, real=0):
    """Runs `iters` optimizer steps; returns (sites, m, v, loss_sum[iters], count[iters])."""
    s = _pts(sites, real).copy()
    t = _pts(targets, real)
    dt = s.dtype
    m = np.zeros_like(s) if m is None else np.ascontiguousarray(m, dt).copy()
    v = np.zeros_like(s) if v is None else np.ascontiguousarray(v, dt).copy()
    lh = np.zeros(iters, np.float64)
    ch = np.zeros(iters, np.int64)
    boxa = None if box is None else np.asarray(box, np.float64)
    mv = None if movable is None else np.ascontiguousarray(movable, np.uint8)
    so = None if slot_of is None else np.ascontiguousarray(slot_of, np.int32)
    rc = lib().vo_fit(real, _p(s), _p(m), _p(v), _p(mv), len(s), _p(t), len(t), _p(so), K, cell, fraction, seed,
                      _p(boxa), t_start, iters, C.byref(sched), _p(lh), _p(ch))
    assert rc == 0
    return s, m, v, lh, ch


def clip(sites, W, H, cell, Kc=32, masks=None, tiles=(0, 0), voxel=0.0, real=0):
    """Clipped Voronoi cells: dict(verts[S,24,2] f32, labels[S,24], count[S], inside[S], unsafe, overflow)."""
    s = _pts(sites, real)
    S = len(s)
    verts = np.zeros((S, KMAX_CELL_VERTS, 2), np.float32)
    labels = np.zeros((S, KMAX_CELL_VERTS), np.int32)
    count = np.zeros(S, np.int32)
    inside = np.zeros(S, np.uint8)
    flags = np.zeros(2, np.int32)
    mk = None if masks is None else np.ascontiguousarray(masks, np.uint64)
    rc = lib().vo_clip(real, _p(s), S, W, H, cell, Kc, _p(mk), tiles[0], tiles[1], voxel, _p(verts), _p(labels),
                       _p(count), _p(inside), _p(flags))
    assert rc == 0
    return dict(verts=verts, labels=labels, count=count, inside=inside, unsafe=int(flags[0]),
                overflow=int(flags[1]))


def boundary(verts, labels, count, inside):
    """Boundary mesh from clipped cells: dict(vertices[V,2], edges[E,2], watertight, loops)."""
    n = len(count)
    verts = np.ascontiguousarray(verts, np.float32)
    labels = np.ascontiguousarray(labels, np.int32)
    count = np.ascontiguousarray(count, np.int32)
    inside = np.ascontiguousarray(inside, np.uint8)
    cap = int(count.sum()) + 8
    ov = np.zeros((cap, 2), np.float32)
    oe = np.zeros((cap, 2), np.int32)
    nv, ne, wt, lp = (np.zeros(1, np.int32) for _ in range(4))
    rc = lib().vo_boundary(n, _p(verts), _p(labels), _p(count), _p(inside), _p(ov), cap, _p(oe), cap, _p(nv),
                           _p(ne), _p(wt), _p(lp))
    assert rc == 0
    return dict(vertices=ov[:nv[0]].copy(), edges=oe[:ne[0]].copy(), watertight=bool(wt[0]), loops=int(lp[0]))


def poisson_dims(W, H, r):
    gw, gh = np.zeros(1, np.int32), np.zeros(1, np.int32)
    cs = np.zeros(1, np.float32)
    lib().vo_poisson_dims(W, H, r, _p(gw), _p(gh), _p(cs))
    return int(gw[0]), int(gh[0]), float(cs[0])


def poisson(W, H, r, trials=30, seed=1, region=None, round_base=0, grid=None):
    """Dart-throwing Poisson disk sampler; returns the (gh*gw, 2) float32 cell grid (x < 0 = empty)."""
    gw, gh, _ = poisson_dims(W, H, r)
    if grid is None:
        grid = np.full((gw * gh, 2), -1.0, np.float32)
    if region is None:
        region = (0, 0, gw, gh)
    rc = lib().vo_poisson(W, H, r, trials, seed, region[0], region[1], region[2], region[3], round_base, _p(grid))
    assert rc == 0
    return grid


def poisson_verify(W, H, r, grid, region=None):
    gw, gh, _ = poisson_dims(W, H, r)
    if region is None:
        region = (0, 0, gw, gh)
    g = np.ascontiguousarray(grid, np.float32)
    return bool(lib().vo_poisson_verify(W, H, r, _p(g), *region))
