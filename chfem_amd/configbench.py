"""Seconds per homogenization of two of the solve path's workloads with the
model resident in HBM, one JSON line:

* c2: the 128^3 two-phase elastic image of random overlapping spheres
  (radius 6-12, inclusion fraction 0.3, seed 2; E 1/10, nu 0.3/0.2), 6 load
  cases, Jacobi-PCG to tol 1e-6, node by node (NBN) and element by element
  (EBE), FP64 and mixed precision;
* c3: the 256^3 permeability model of solid spheres (radius 6-12) until
  porosity 0.35 (seed 4), 3 load cases, Jacobi-MINRES to tol 1e-6, FP64 and
  mixed precision.

Each run prepares the model, solves load case 0 once untimed (the autotuner
and the solver's buffers), then times every load case in turn
(hmgSolveLoadCase) and reports the iterations, refinement sweeps and the
effective-tensor columns.

    python -m chfem_amd.configbench --config c3
"""
import argparse
import ctypes as C
import json
import os
import shutil
import tempfile
import time

import numpy as np

from . import io as cio
from . import synthetic
from .lib import SolveStats, lib, ptr

PRECISIONS = {"fp64": 0, "mixed": 1}
STRATEGIES = {"nbn": 0, "ebe": 1}


def run(L, nf, flat, loads, solver, strategy, precision, tol):
    assert L.hmgInit(nf.encode(), None, None, ptr(flat)) == 1
    try:
        L.hmgSetSolverFlag(solver)
        L.hmgSetParallelStrategyFlag(strategy)
        L.hmgSetPrecisionFlag(precision)
        L.hmgSetSolverTolerance(tol)
        assert L.hmgPrepare() == 1
        ncol = 6 if loads == 6 else 3
        col = np.zeros(ncol)
        L.hmgSolveLoadCase(0, None, None, col.ctypes.data_as(C.c_void_p))  # untimed: autotuner, buffers
        cols, its, sweeps, secs, ok = [], [], [], [], []
        st = SolveStats()
        for d in range(loads):
            col = np.zeros(ncol)
            t = time.perf_counter()
            conv = L.hmgSolveLoadCase(d, None, None, col.ctypes.data_as(C.c_void_p))
            secs.append(time.perf_counter() - t)
            L.cudapcgGetSolveStats(C.byref(st))
            cols.append(col.tolist())
            its.append(int(st.iterations))
            sweeps.append(int(st.sweeps))
            ok.append(bool(conv))
        L.hmgRelease()
        return {"seconds": sum(secs), "seconds_per_load": secs, "iterations": its, "sweeps": sweeps,
                "converged": ok, "C_columns": cols}
    finally:
        L.hmgEnd()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=["c2", "c3"], default="c3")
    ap.add_argument("--size", type=int, default=0, help="edge length (default 128 for c2, 256 for c3)")
    ap.add_argument("--tol", type=float, default=1e-6)
    a = ap.parse_args()
    L = lib()
    tmp = tempfile.mkdtemp(prefix="chfem_cfg_")
    try:
        f = os.path.join(tmp, "model.raw")
        if a.config == "c2":
            n = a.size or 128
            img = synthetic.spheres((n, n, n), rmin=6, rmax=12, fraction=0.3, seed=2)
            cio.export_for_chfem(f, img, 1, {0: (1.0, 0.3), 255: (10.0, 0.2)}, export_raw=False,
                                 solver_tolerance=a.tol, solver_maxiter=20000)
            out = {"config": "c2", "workload": f"elastic_{n}^3_spheres_r6-12_f0.3_seed2_E1-10_nu0.3-0.2_6loads",
                   "tol": a.tol, "runs": {}}
            flat = np.ascontiguousarray(img).reshape(-1)
            for sname, sflag in STRATEGIES.items():
                for pname, pflag in PRECISIONS.items():
                    r = run(L, f[:-4] + ".nf", flat, 6, 0, sflag, pflag, a.tol)
                    out["runs"][f"{sname}_{pname}"] = r
                    print(f"c2 {sname} {pname}: {r['seconds']:.3f} s, iterations {r['iterations']}", flush=True)
        else:
            n = a.size or 256
            img, seed = synthetic.porous_spheres((n, n, n), rmin=6, rmax=12, porosity=0.35, seed=4)
            cio.export_for_chfem(f, img, 2, None, export_raw=False, solver_tolerance=a.tol, solver_maxiter=20000)
            out = {"config": "c3", "workload": f"stokes_{n}^3_solid_spheres_r6-12_porosity0.35_seed{seed}_3loads",
                   "tol": a.tol, "porosity": float((img == 0).mean()), "runs": {}}
            flat = np.ascontiguousarray(img).reshape(-1)
            for pname, pflag in PRECISIONS.items():
                r = run(L, f[:-4] + ".nf", flat, 3, 1, 0, pflag, a.tol)
                out["runs"][f"minres_{pname}"] = r
                print(f"c3 minres {pname}: {r['seconds']:.3f} s, iterations {r['iterations']}, "
                      f"sweeps {r['sweeps']}", flush=True)
        print(json.dumps(out))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
