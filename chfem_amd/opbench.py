"""Operator micro-benchmark: q = A d for a seeded two-phase sphere image, and a
plain copy kernel of known traffic for scaling hardware byte counters.

    python -m chfem_amd.opbench --physics elastic --size 256 --precision mixed --reps 20

Thermal and elastic use bench.py's sphere image (seed 3, inclusion fraction
0.3); fluid uses solid spheres until the porosity is 0.35 (--porosity; seed
4, bumped until the pore space percolates), the Stokes model of the
permeability workload.
"""
import argparse
import json
import os
import tempfile

import numpy as np

from . import io as cio
from . import synthetic
from .lib import lib, ptr

PHYS = {"thermal": 0, "elastic": 1, "fluid": 2}
BYTES = {("thermal", 0): 17.0, ("thermal", 1): 9.0, ("elastic", 0): 49.0, ("elastic", 1): 25.0}
# CG-loop form with the direction update inside the operator: image, z and
# the previous direction read, the direction and q written (the iterate is
# updated per window of iterations outside the operator)
BYTES_DIR = {("thermal", 0): 1.0 + 4 * 8.0, ("thermal", 1): 1.0 + 4 * 4.0,
             ("elastic", 0): 1.0 + 4 * 24.0, ("elastic", 1): 1.0 + 4 * 12.0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--physics", default="elastic", choices=list(PHYS))
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--strategy", type=int, default=0)
    ap.add_argument("--copy", action="store_true", help="also time the calibration copy kernel")
    ap.add_argument("--dot", action="store_true", help="with the fused v.Av reduction, as inside the solver")
    ap.add_argument("--dir", action="store_true",
                    help="the CG loop's form: direction update d = z + beta d inside the operator (Walsh NBN)")
    ap.add_argument("--both", action="store_true", help="time the --dot form, then the --dir form")
    ap.add_argument("--porosity", type=float, default=0.35, help="fluid: porosity of the sphere packing")
    ap.add_argument("--cg1", action="store_true",
                    help="also time the single-kernel CG form (z = d_old - beta' d_old2 - alpha M^-1 q inside the "
                         "operator too)")
    a = ap.parse_args()
    n = a.size
    if a.physics == "fluid":
        # solid spheres (radius 6-12) until porosity <= 0.35, pore space percolating
        img, _ = synthetic.porous_spheres((n, n, n), rmin=6, rmax=12, porosity=a.porosity, seed=4)
        props = None
    else:
        img = synthetic.spheres((n, n, n), rmin=6, rmax=12, fraction=0.3, seed=3)
        props = {0: 1.0, 255: 10.0} if a.physics == "thermal" else {0: (1.0, 0.3), 255: (10.0, 0.2)}
    d = tempfile.mkdtemp(prefix="chfem_op_")
    f = os.path.join(d, "img.raw")
    cio.export_for_chfem(f, img, PHYS[a.physics], props, export_raw=False)
    L = lib()
    assert L.hmgInit((f[:-4] + ".nf").encode(), None, None, ptr(np.ascontiguousarray(img).reshape(-1))) == 1
    prec = 1 if a.precision == "mixed" else 0
    L.hmgSetParallelStrategyFlag(a.strategy)
    L.hmgSetPrecisionFlag(prec)
    assert L.hmgPrepare() == 1
    N = n ** 3
    out = {"physics": a.physics, "size": n, "precision": a.precision, "strategy": a.strategy}
    if a.physics == "fluid":
        # Stokes: pore nodes read a DOF id (4 B) and 4 values, write 4; border
        # nodes read their DOF id and fluid key (5 B) and 1 value, write 1
        import ctypes as C
        ne, nd_, nv, nb = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
        L.hmgGetSizes(C.byref(ne), C.byref(nd_), C.byref(nv), C.byref(nb))
        w = 4 if prec == 1 else 8
        fluid_bytes = nv.value * (4 + 8 * w) + nb.value * (5 + 2 * w)
        out.update({"pore_nodes": nv.value, "border_nodes": nb.value, "dofs": nd_.value})
    forms = ["dot", "dir"] if a.both else ["dir" if a.dir else ("dot" if a.dot else "plain")]
    if a.cg1:
        forms.append("cg1")
    for form in forms:
        ms = L.cudapcgOperatorBench(prec, a.reps, {"plain": 0, "dot": 1, "dir": 2, "cg1": 3}[form])
        if ms < 0:
            raise SystemExit("operator bench failed (the direction form needs the Walsh-basis NBN operator)")
        if a.physics == "fluid":
            nbytes = fluid_bytes
        else:
            w = 1 if a.physics == "thermal" else 3
            vb = (4 if prec == 1 else 8) * w
            # cg1: d_old2, qt_old, d_old read + d, qt written (z recovered
            # from the two previous directions)
            per = {"dir": BYTES_DIR[(a.physics, prec)], "cg1": 1.0 + 5 * vb}.get(form, BYTES[(a.physics, prec)])
            nbytes = per * N
        out[form] = {"avg_launch_ms": ms, "algorithmic_bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9}
    L.hmgRelease()
    L.hmgEnd()
    if a.copy:
        nc = 64 * 1024 * 1024
        cms = L.chfemCopyKernelBench(nc, a.reps)
        out["copy"] = {"entries": nc, "bytes_read": 8 * nc, "bytes_written": 8 * nc, "avg_launch_ms": cms,
                       "GBps": 16 * nc / (cms * 1e-3) / 1e9}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
