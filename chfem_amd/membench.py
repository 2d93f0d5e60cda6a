"""Device memory of the permeability solvers: the Stokes model of solid
spheres (radius 6-12) until porosity 0.35 at --size^3, solved with MINRES
(x, r, d, q, s plus the stored diagonal and the work buffers of the loop),
MINRES3 (x, z, d) and MINRES2 (z, d, and the reduced velocity: every velocity
DOF in the FULL reduction mode, one component in ONLYDIR), FP64, for a
bounded number of iterations; device memory is read with hipMemGetInfo
before the model is prepared and after its first solve.

    python -m chfem_amd.membench --size 512 --iterations 40
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

# name: (solver flag, xreduce mode 2 FULL / 1 ONLYDIR)
SOLVERS = {"MINRES": (1, 2), "MINRES3": (3, 2), "MINRES2": (5, 2), "MINRES2_ONLYDIR": (5, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--iterations", type=int, default=40)
    a = ap.parse_args()
    n = a.size
    t0 = time.perf_counter()
    img, seed = synthetic.porous_spheres((n, n, n), rmin=6, rmax=12, porosity=0.35, seed=4)
    t_img = time.perf_counter() - t0
    d = tempfile.mkdtemp(prefix="chfem_mem_")
    f = os.path.join(d, "pores.raw")
    cio.export_for_chfem(f, img, 2, None, export_raw=False, solver_tolerance=1e-12, solver_maxiter=a.iterations)
    L = lib()
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]

    def free_bytes():
        fr, tot = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(fr), C.byref(tot)) == 0
        return fr.value, tot.value

    flat = np.ascontiguousarray(img).reshape(-1)
    out = {"size": n, "seed": seed, "porosity": float((img == 0).mean()), "image_seconds": t_img,
           "iterations": a.iterations, "solvers": {}}
    del img
    for name, (flag, xmode) in SOLVERS.items():
        assert L.hmgInit((f[:-4] + ".nf").encode(), None, None, ptr(flat)) == 1
        try:
            ne, nd, nv, nb = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
            L.hmgGetSizes(C.byref(ne), C.byref(nd), C.byref(nv), C.byref(nb))
            L.hmgSetSolverFlag(flag)
            L.hmgSetXReduceFlag(xmode)
            L.hmgSetMaxIterationsOverride(a.iterations)
            L.hmgSetSolverTolerance(1e-12)
            before, total = free_bytes()
            assert L.hmgPrepare() == 1
            prepared, _ = free_bytes()
            col = np.zeros(3)
            t1 = time.perf_counter()
            L.hmgSolveLoadCase(0, None, None, col.ctypes.data_as(C.c_void_p))
            t2 = time.perf_counter()
            solved, _ = free_bytes()
            st = SolveStats()
            L.cudapcgGetSolveStats(C.byref(st))
            L.hmgRelease()
            vec = 8.0 * nd.value
            out["solvers"][name] = {
                "device_GB_model_and_solver": (before - solved) / 1e9,
                "device_GB_at_prepare": (before - prepared) / 1e9,
                "fp64_vectors_equivalent": (before - solved) / vec,
                "iterations": st.iterations, "seconds": t2 - t1,
                "ms_per_iteration": 1e3 * (t2 - t1) / max(st.iterations, 1)}
            out.update({"dofs": nd.value, "pore_nodes": nv.value, "border_nodes": nb.value,
                        "fp64_vector_GB": vec / 1e9, "device_total_GB": total / 1e9})
            print(f"{name}: {out['solvers'][name]}", flush=True)
        finally:
            L.hmgEnd()
    print(json.dumps(out), flush=True)
    shutil.rmtree(d)


if __name__ == "__main__":
    main()
