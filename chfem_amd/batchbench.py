"""Batched-solver benchmark: many independent samples of uneven sizes, half
thermal and half elastic, every load case solved together on the GPU
(``chfemBatch*``), against the same samples solved one at a time by the
single-system solver in the same precision.

    python -m chfem_amd.batchbench --samples 64 --min 48 --max 160 --tol 1e-8

Sample s (seed 100 + s) has edges drawn uniformly from [min, max] per axis.
Even samples: conductivity of a thresholded Gaussian random field (volume
fraction 0.4, k 1 and 50).  Odd samples: elasticity of overlapping spheres
(radius 6-12, inclusion fraction 0.3, E 1 / 20, nu 0.3 / 0.2).  Both
methods run mixed precision (FP32 CG, FP64 scalars and residual
replacements) to the same tolerance.  Prints one JSON line with the device
seconds of each method (RHS to convergence, summed over load cases for the
one-at-a-time solves), wall seconds, iteration totals and the largest
relative difference between the two methods' effective matrices.
"""
import argparse
import ctypes as C
import json
import os
import sys
import tempfile
import time

import numpy as np

from . import batch, synthetic
from . import io as cio
from .lib import SolveStats, lib, ptr


def make_samples(n, lo, hi, seed0=100):
    out = []
    for s in range(n):
        rng = np.random.default_rng(seed0 + s)
        shape = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        if s % 2 == 0:
            img = synthetic.grf(shape, sigma=2.0, vf=0.4, seed=seed0 + s)
            out.append(("conductivity", img, {0: 1.0, 255: 50.0}))
        else:
            img = synthetic.spheres(shape, rmin=6, rmax=12, fraction=0.3, seed=seed0 + s)
            out.append(("elasticity", img, {0: (1.0, 0.3), 255: (20.0, 0.2)}))
    return out


def solve_one(img, prop, mp, tol, maxit):
    """One sample through the single-system solver (mixed precision, all load
    cases, model resident): (C, device seconds of its solves, iterations)."""
    L = lib()
    ph = 0 if prop == "conductivity" else 1
    d = tempfile.mkdtemp(prefix="chfem_bb_")
    nf = os.path.join(d, "img.raw")
    cio.export_for_chfem(nf, img, ph, mp, voxel_size=1.0, export_raw=False, solver_tolerance=tol,
                         solver_maxiter=maxit)
    nf = nf[:-4] + ".nf"
    flat = np.ascontiguousarray(img).reshape(-1)
    try:
        if L.hmgInit(nf.encode(), None, None, ptr(flat)) != 1:
            raise RuntimeError("hmgInit failed")
        L.hmgSetPrecisionFlag(1)
        L.hmgSetHomogenizationFlag(6)
        Cm = np.zeros(36)
        if L.hmgPrepare() != 1:
            raise RuntimeError("hmgPrepare failed")
        L.hmgSolveResident(ptr(Cm))
        st = (SolveStats * 6)()
        cnt = C.c_uint(0)
        L.hmgGetSolveStats(st, C.byref(cnt))
        L.hmgRelease()
        L.hmgSetPrecisionFlag(0)
        L.hmgEnd()
    finally:
        os.remove(nf)
        os.rmdir(d)
    n = 6 if ph == 1 else 3
    secs = sum(st[i].solve_seconds for i in range(cnt.value))
    its = sum(st[i].iterations for i in range(cnt.value))
    return Cm[:n * n].reshape(n, n), secs, its


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=64)
    ap.add_argument("--min", type=int, default=48)
    ap.add_argument("--max", type=int, default=160)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--maxit", type=int, default=20000)
    a = ap.parse_args()
    t0 = time.perf_counter()
    samples = make_samples(a.samples, a.min, a.max)
    gen_s = time.perf_counter() - t0
    print(f"[batchbench] {a.samples} samples generated in {gen_s:.1f} s", file=sys.stderr, flush=True)
    out = {"samples": a.samples, "edges": [a.min, a.max], "tol": a.tol,
           "voxels": int(sum(img.size for _, img, _ in samples))}
    Cb = {}
    for prop in ("conductivity", "elasticity"):
        idx = [k for k, (p, _, _) in enumerate(samples) if p == prop]
        if not idx:
            continue
        t1 = time.perf_counter()
        Cs, info = batch.batch_homogenize([(samples[k][1], samples[k][2]) for k in idx], prop, tol=a.tol,
                                          maxit=a.maxit)
        wall = time.perf_counter() - t1
        for k, Cm in zip(idx, Cs):
            Cb[k] = Cm
        its = int(sum(s[0] for st in info["stats"] for s in st))
        out[f"batched_{prop}"] = {"samples": len(idx), "device_s": info["seconds"], "wall_s": wall,
                                  "converged": info["converged"], "iterations": its}
        print(f"[batchbench] batched {prop}: {len(idx)} samples, {info['seconds']:.3f} s device, "
              f"{wall:.1f} s wall", file=sys.stderr, flush=True)
    seq = {"conductivity": [0.0, 0.0, 0], "elasticity": [0.0, 0.0, 0]}
    diff = 0.0
    for k, (prop, img, mp) in enumerate(samples):
        t1 = time.perf_counter()
        Cs, secs, its = solve_one(img, prop, mp, a.tol, a.maxit)
        seq[prop][0] += secs
        seq[prop][1] += time.perf_counter() - t1
        seq[prop][2] += its
        diff = max(diff, float(np.abs(Cs - Cb[k]).max() / np.abs(Cs).max()))
    for prop, (secs, wall, its) in seq.items():
        if f"batched_{prop}" in out:
            out[f"sequential_{prop}"] = {"device_s": secs, "wall_s": wall, "iterations": its}
    dev_b = sum(out[k]["device_s"] for k in out if k.startswith("batched_"))
    dev_s = sum(out[k]["device_s"] for k in out if k.startswith("sequential_"))
    out["device_s_batched"] = dev_b
    out["device_s_sequential"] = dev_s
    out["speedup_device"] = dev_s / dev_b if dev_b > 0 else None
    out["max_rel_diff_C"] = diff
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
