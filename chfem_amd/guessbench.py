"""Learned initial guess at scale: a network trained on many small seeded
elastic images, applied to a large one.

    python -m chfem_amd.guessbench --train-samples 2000 --train-size 64 --size 512

Training data: --train-samples seeded 64^3 two-phase images, half thresholded
Gaussian random fields (volume fraction 0.4), half overlapping spheres (radius
6-12, inclusion fraction 0.3); matrix E 1, nu 0.3, inclusion E 10, nu 0.2; the
6 load cases of each solved by the library (mixed precision, tol 1e-6).  The
network is the library's convolutional guess (Adam, --epochs passes).

Test: a --size^3 sphere image (seed 5) of the same recipe, all 6 load cases
solved by mixed-precision PCG to --tol twice, from the zero guess and from
the learned guess, with the model resident.  Prints one JSON line: device
seconds and iterations of both, the initial relative residuals of the guess,
and the largest relative difference between the two effective matrices.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import guess, synthetic
from . import io as cio
from .lib import SolveStats, lib, ptr

MP = {0: (1.0, 0.3), 255: (10.0, 0.2)}


def log(m):
    print(m, file=sys.stderr, flush=True)


def training_image(k, size, seed0=2000):
    if k % 2 == 0:
        return synthetic.grf((size, size, size), sigma=2.0, vf=0.4, seed=seed0 + k)
    return synthetic.spheres((size, size, size), rmin=6, rmax=12, fraction=0.3, seed=seed0 + k)


def solve_resident(img, tol, net):
    """All 6 load cases of img, mixed precision, model resident; returns
    (C, per-load stats, wall seconds of the solve)."""
    L = lib()
    d = tempfile.mkdtemp(prefix="chfem_gb_")
    nf = os.path.join(d, "img.raw")
    cio.export_for_chfem(nf, img, 1, MP, voxel_size=1.0, export_raw=False, solver_tolerance=tol,
                         solver_maxiter=40000)
    nf = nf[:-4] + ".nf"
    flat = np.ascontiguousarray(img).reshape(-1)
    if L.hmgInit(nf.encode(), None, None, ptr(flat)) != 1:
        raise RuntimeError("hmgInit failed")
    try:
        L.hmgSetPrecisionFlag(1)
        L.hmgSetHomogenizationFlag(6)
        L.hmgSetInitialGuessNetwork(net.h if net is not None else None)
        Cm = np.zeros(36)
        if L.hmgPrepare() != 1:
            raise RuntimeError("hmgPrepare failed")
        t0 = time.perf_counter()
        ok = L.hmgSolveResident(ptr(Cm))
        wall = time.perf_counter() - t0
        st = (SolveStats * 6)()
        cnt = C.c_uint(0)
        L.hmgGetSolveStats(st, C.byref(cnt))
        L.hmgRelease()
        if ok != 1:
            raise RuntimeError("hmgSolveResident failed")
    finally:
        L.hmgSetInitialGuessNetwork(None)
        L.hmgSetPrecisionFlag(0)
        L.hmgEnd()
        os.remove(nf)
        os.rmdir(d)
    stats = [{"iterations": st[i].iterations, "seconds": st[i].solve_seconds, "x0_residual": st[i].x0_residual,
              "x0_used": st[i].x0_used} for i in range(cnt.value)]
    return Cm.reshape(6, 6), stats, wall


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train-samples", type=int, default=2000)
    ap.add_argument("--train-size", type=int, default=64)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--tol", type=float, default=1e-6)
    a = ap.parse_args()
    out = {"train_samples": a.train_samples, "train_size": a.train_size, "epochs": a.epochs, "size": a.size,
           "tol": a.tol}
    t0 = time.perf_counter()
    samples = []
    for k in range(a.train_samples):
        samples.append(guess.solve_fluctuations(training_image(k, a.train_size), MP, "elastic", tol=1e-6))
        if (k + 1) % 100 == 0 or k + 1 == a.train_samples:
            log(f"[guessbench] training data {k + 1}/{a.train_samples} ({time.perf_counter() - t0:.0f} s)")
    out["data_seconds"] = time.perf_counter() - t0
    net = guess.LearnedGuess("elastic", seed=0)
    t1 = time.perf_counter()
    hist = net.train(samples, epochs=a.epochs, lr=2e-3, batch=4, log=lambda m: log(f"[guessbench] {m}"))
    out["train_seconds"] = time.perf_counter() - t1
    out["loss"] = [hist[0], hist[-1]]
    del samples
    img = synthetic.spheres((a.size,) * 3, rmin=6, rmax=12, fraction=0.3, seed=5)
    log(f"[guessbench] {a.size}^3 image ready; zero-guess solve")
    C0, s0, w0 = solve_resident(img, a.tol, None)
    log(f"[guessbench] zero guess: {sum(s['iterations'] for s in s0)} iterations, {w0:.1f} s; learned-guess solve")
    C1, s1, w1 = solve_resident(img, a.tol, net)
    log(f"[guessbench] learned guess: {sum(s['iterations'] for s in s1)} iterations, {w1:.1f} s")
    out["zero_guess"] = {"seconds": w0, "solve_seconds": sum(s["seconds"] for s in s0),
                         "iterations": [s["iterations"] for s in s0]}
    out["learned_guess"] = {"seconds": w1, "solve_seconds": sum(s["seconds"] for s in s1),
                            "iterations": [s["iterations"] for s in s1],
                            "x0_residual": [s["x0_residual"] for s in s1], "x0_used": [s["x0_used"] for s in s1]}
    out["speedup"] = w0 / w1 if w1 > 0 else None
    out["max_rel_diff_C"] = float(np.abs(C1 - C0).max() / np.abs(C0).max())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
