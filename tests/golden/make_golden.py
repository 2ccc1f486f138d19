"""Writes the golden fixtures under tests/golden/.

kat.json holds known-answer values computed here in plain Python integers and
fractions, independently of the oracle: the PCG hash of 0..3, the two rand01
draws, and the VoroLoss hand case (c0=(0,0), cj=(1,0), p=(0.25,0.3)).  The
hash values agree with the ones listed for the reference's pcg_hash /
rand01 (src/voromesh/voro_optimizer.cu:69-78, src/poisson/poisson_disk.cu:47-58).

config0_small.npz is a small seeded case (Poisson sites r=0.05 seed 0, 1,024
star targets) with the oracle's k-NN lists, per-target terms and gradient, as a
regression anchor for the GPU parity tests.

Run:  python tests/golden/make_golden.py
"""
import json
import os
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def pcg_hash(v):
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word) & 0xFFFFFFFF


def rand01_loss(index, it, seed):
    h = pcg_hash(index ^ pcg_hash(((it * 2654435761) & 0xFFFFFFFF) ^ pcg_hash(seed)))
    return (h & 0xFFFFFF) / 16777216.0


def rand01_poisson(cell, rnd, seed, stream):
    h = pcg_hash(cell ^ pcg_hash(((rnd * 2654435761) & 0xFFFFFFFF) ^
                                 pcg_hash((seed + stream * 0x9E3779B9) & 0xFFFFFFFF)))
    return (h & 0xFFFFFF) / 16777216.0


def loss_hand():
    # d = <u, e>/|e| - |e|/2 with u = p - c0, e = cj - c0; L = d^2;
    # dd/dcj = (u - <u,n> n)/l - n/2, dd/dc0 = -n - dd/dcj, dL/dc = 2 d dd/dc.
    c0, cj, p = (Fraction(0), Fraction(0)), (Fraction(1), Fraction(0)), (Fraction(1, 4), Fraction(3, 10))
    u = (p[0] - c0[0], p[1] - c0[1])
    e = (cj[0] - c0[0], cj[1] - c0[1])
    l = Fraction(1)  # |e| = 1
    n = (e[0] / l, e[1] / l)
    un = u[0] * n[0] + u[1] * n[1]
    d = un - l / 2
    dcj = ((u[0] - un * n[0]) / l - n[0] / 2, (u[1] - un * n[1]) / l - n[1] / 2)
    dc0 = (-n[0] - dcj[0], -n[1] - dcj[1])
    g0 = [float(2 * d * dc0[0]), float(2 * d * dc0[1])]
    gj = [float(2 * d * dcj[0]), float(2 * d * dcj[1])]
    return {"d": float(d), "loss": float(d * d), "grad": [g0, gj]}


def main():
    kat = {
        "pcg_hash_0_3": [f"0x{pcg_hash(i):08x}" for i in range(4)],
        "rand01_loss_i0_3_iter1_seed1": [rand01_loss(i, 1, 1) for i in range(4)],
        "rand01_poisson_c0_r0_s1": [rand01_poisson(0, 0, 1, 0), rand01_poisson(0, 0, 1, 1)],
        "loss_hand": loss_hand(),
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    print(json.dumps(kat))

    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "bb-temp-testing_amd"))
    import numpy as np
    import oracle as O
    from vmhip import workloads as wl
    grid = O.poisson(1.0, 1.0, 0.05, 30, 0)
    sites = wl.live_sites(grid)
    targets = wl.star(1024, seed=0)
    idx, d2 = O.knn(sites, targets, 16, 1.0, 1.0, 0.1, brute=True)
    r = O.loss_grad(sites, targets, 16, 1.0, 1.0, 0.1, per_target=True)
    np.savez_compressed(os.path.join(HERE, "config0_small.npz"), grid=grid, sites=sites, targets=targets,
                        knn=idx, knn_d2=d2, c0=r["c0"], bestj=r["bestj"], loss_t=r["loss_t"], grad=r["grad"])


if __name__ == "__main__":
    main()
