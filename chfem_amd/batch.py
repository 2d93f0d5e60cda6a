"""Batched homogenization: every load case of many samples (one analysis
type, any sizes) solved together on the GPU (``chfemBatch*`` in
include/chfem_amd.h).  The reference solves one model and one load case at a
time (chfem/src/femhmg/femhmg.c:848-1198)."""
import ctypes as C
import os
import tempfile

import numpy as np

from . import io as cio
from .lib import SolveStats, lib, ptr

PROPERTIES = {"conductivity": 0, "elasticity": 1}


def batch_homogenize(samples, property, tol=1e-8, maxit=20000, voxel_size=1.0, guess=None):
    """samples: list of (uint8 image (z, y, x), mat_props) -> list of effective
    matrices, and a dict with per-sample, per-load-case statistics and the
    batch's device time.  guess: a trained chfem_amd.guess.LearnedGuess of
    the same physics, whose prediction starts every load case (statistics
    then carry x0_used and x0_residual)."""
    if property not in PROPERTIES:
        raise ValueError(f"Invalid property: {property}")
    analysis = PROPERTIES[property]
    L = lib()
    h = L.chfemBatchCreate()
    d = tempfile.mkdtemp(prefix="chfem_batch_")
    files = []
    try:
        for k, (img, mp) in enumerate(samples):
            if img.dtype != np.uint8:
                raise ValueError("Domain must be uint8 dtype.")
            f = os.path.join(d, f"s{k}.raw")
            cio.export_for_chfem(f, img, analysis, mp, voxel_size=voxel_size, export_raw=False,
                                 solver_tolerance=tol, solver_maxiter=maxit)
            nf = f[:-4] + ".nf"
            files.append(nf)
            flat = np.ascontiguousarray(img).reshape(-1)
            if L.chfemBatchAddSample(h, nf.encode(), ptr(flat)) != 1:
                raise RuntimeError(f"chfemBatchAddSample failed for sample {k}")
        if guess is not None and L.chfemBatchSetInitialGuessNetwork(h, guess.h) != 1:
            raise RuntimeError("chfemBatchSetInitialGuessNetwork failed")
        rc = L.chfemBatchSolve(h, tol, maxit)
        if rc == 0:
            raise RuntimeError("chfemBatchSolve failed")
        n = 6 if analysis == 1 else 3
        out, stats = [], []
        for k in range(len(samples)):
            Cm = np.zeros(n * n)
            L.chfemBatchGetC(h, k, ptr(Cm))
            out.append(Cm.reshape(n, n))
            st = (SolveStats * 6)()
            cnt = C.c_uint(0)
            L.chfemBatchGetStats(h, k, st, C.byref(cnt))
            stats.append([(st[i].iterations, st[i].sweeps, st[i].residual, st[i].x0_used, st[i].x0_residual)
                          for i in range(cnt.value)])
        info = {"converged": rc == 1, "seconds": L.chfemBatchSeconds(h), "stats": stats}
        return out, info
    finally:
        L.chfemBatchDestroy(h)
        for nf in files:
            if os.path.exists(nf):
                os.remove(nf)
        os.rmdir(d)
