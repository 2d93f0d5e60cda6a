"""Shared test helpers: oracle models of seeded images, and a cudapcg solver
set up through the C ABI the way a reference caller does
(chfem/src/femhmg/femhmg.c:914-1002)."""
import ctypes as C

import numpy as np

from chfem_amd import synthetic
from oracle import oracle as O

THERMAL, ELASTIC, FLUID = 0, 1, 2
PCG_FLAG = {THERMAL: 1, ELASTIC: 3, FLUID: 5}


def model(analysis, img, props, colors=(0, 255), voxel=1.0):
    """Oracle model of an image (z, y, x) whose colours are listed in .nf order."""
    nz, ny, nx = img.shape
    em = synthetic.elem_map(img, list(colors))
    return O.Model(analysis, (nx, ny, nz), em, props, voxel)


def reference_local(m):
    """Local matrices in the reference layout the C ABI takes
    (thermal: 40 stored entries per material)."""
    K, _ = m.local()
    if m.analysis == THERMAL:
        out = np.zeros(40 * m.nmat)
        for t in range(m.nmat):
            O.lib().orc_thermal_pack40(O._p(np.ascontiguousarray(K[64 * t:64 * t + 64])), 1.0,
                                       O._p(out[40 * t:]))
        return out
    return K


class GpuSolver:
    """cudapcg solver over an oracle model, through include/cudapcg.h."""

    def __init__(self, m, strategy=0, solver=0, precond=1, precision=0, tol=1e-6, maxit=10000):
        from chfem_amd.lib import CudapcgModel, lib, ptr
        self.L = lib()
        self.m = m
        L = self.L
        assert L.cudapcgInit(PCG_FLAG[m.analysis], strategy) == 1
        p = L.cudapcgNewModel()
        cm = C.cast(p, C.POINTER(CudapcgModel)).contents
        cm.parStrategy_flag = strategy
        cm.poremap_flag = 1
        cm.nrows, cm.ncols, cm.nlayers = m.ny, m.nx, m.nz
        cm.nelem = m.N
        cm.nvars = m.ndof
        cm.nvarspernode = 3
        # (femhmg.c:82: the velocity DOFs for Stokes, every DOF otherwise)
        cm.nhmgvars = 3 * m.nvel if m.analysis == FLUID else m.ndof
        cm.nkeys = 1 if m.analysis == FLUID else m.nmat
        cm.localmtxdim = {THERMAL: 40, ELASTIC: 576, FLUID: 1024}[m.analysis]
        if m.analysis == FLUID:
            cm.nporenodes = m.nvel
            cm.nbordernodes = m.nborder
        assert L.cudapcgSetModel(p) == 1
        if m.analysis == FLUID:
            fk, ids, d2n, bk = m.fluid_maps()
            self._keep = (ids, d2n, bk)
            assert L.cudapcgSetPeriodic2DOFMap(ptr(ids)) == 1
            assert L.cudapcgSetPoreMap(ptr(np.ascontiguousarray(bk)) if bk.size else ptr(np.zeros(1, np.uint8))) == 1
            assert L.cudapcgSetDOF2PeriodicMap(ptr(d2n)) == 1
        else:
            img = m.elem.copy() if strategy == 1 else m.keys()
            assert L.cudapcgSetImage(ptr(img)) == 1
        Kref = reference_local(m)
        assert L.cudapcgSetLclMtxs(ptr(Kref)) == 1
        L.cudapcgSetSolver(solver)
        L.cudapcgSetPreconditioner(precond)
        L.cudapcgSetPrecision(precision)
        L.cudapcgSetNumTol(tol)
        L.cudapcgSetMaxIterations(maxit)
        self.ptr = ptr

    def apply(self, v):
        q = np.zeros(self.m.ndof)
        assert self.L.cudapcgApplyOperator(self.ptr(np.ascontiguousarray(v, dtype=np.float64)), self.ptr(q)) == 1
        return q

    def apply32(self, v):
        q = np.zeros(self.m.ndof, dtype=np.float32)
        assert self.L.cudapcgApplyOperatorF32(self.ptr(np.ascontiguousarray(v, dtype=np.float32)), self.ptr(q)) == 1
        return q

    def precond(self, v):
        z = np.zeros(self.m.ndof)
        assert self.L.cudapcgApplyPreconditioner(self.ptr(np.ascontiguousarray(v, dtype=np.float64)),
                                                 self.ptr(z)) == 1
        return z

    def solve(self, b, x0=None):
        from chfem_amd.lib import SolveStats
        b = np.ascontiguousarray(b, dtype=np.float64)
        assert self.L.cudapcgSetRHS(self.ptr(b)) == 1
        if x0 is not None:
            self._x0 = np.ascontiguousarray(x0, dtype=np.float64)
            assert self.L.cudapcgSetX0(self.ptr(self._x0), 0) == 1
        x = np.zeros(self.m.ndof)
        ok = self.L.cudapcgSolve(self.ptr(x))
        st = SolveStats()
        self.L.cudapcgGetSolveStats(C.byref(st))
        return x, ok, st

    def close(self):
        self.L.cudapcgEnd()


class Resident:
    """A model loaded through the reference's entry points the way
    chfem.wrapper.run does (a .nf written by chfem_amd.io and a NumPy image
    through hmgInit), with the solver flags set, then kept resident on the GPU
    (hmgPrepare) so load cases can be solved one by one (hmgSolveLoadCase)."""

    def __init__(self, img, analysis, mp, tol, precision=0, strategy=0, solver=None, maxit=20000, voxel=1.0):
        import os
        import tempfile
        from chfem_amd import io as cio
        from chfem_amd.lib import lib
        self.L = lib()
        self.analysis = analysis
        self.dim = 6 if analysis == ELASTIC else 3
        self._dir = tempfile.mkdtemp(prefix="chfem_res_")
        nf = os.path.join(self._dir, "model.raw")
        cio.export_for_chfem(nf, img, analysis, mp, voxel_size=voxel, export_raw=False, solver_tolerance=tol,
                             solver_maxiter=maxit)
        self.nf = nf[:-4] + ".nf"
        self._flat = np.ascontiguousarray(img, dtype=np.uint8).reshape(-1)
        L = self.L
        assert L.hmgInit(self.nf.encode(), None, None, self._flat.ctypes.data_as(C.c_void_p)) == 1
        self.open = True
        L.hmgSetParallelStrategyFlag(strategy)
        L.hmgSetSolverFlag(solver if solver is not None else (1 if analysis == FLUID else 0))
        L.hmgSetPreConditionerFlag(1)
        L.hmgSetPrecisionFlag(precision)
        L.hmgSetSolverTolerance(tol)
        L.hmgSetMaxIterationsOverride(maxit)
        self.prepared = False

    def sizes(self):
        ne, nd, nv, nb = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
        self.L.hmgGetSizes(C.byref(ne), C.byref(nd), C.byref(nv), C.byref(nb))
        return ne.value, nd.value, nv.value, nb.value

    def maps(self):
        ne, nd, nv, nb = self.sizes()
        keys = np.zeros(ne, np.uint16)
        fk = np.zeros(ne, np.uint8)
        ids = np.zeros(ne, np.uint32)
        d2n = np.zeros(max(nv + nb, 1), np.uint32)
        bk = np.zeros(max(nb, 1), np.uint8)
        assert self.L.hmgGetMaps(*(a.ctypes.data_as(C.c_void_p) for a in (keys, fk, ids, d2n, bk))) == 1
        return keys, fk, ids, d2n[:nv + nb], bk[:nb]

    def solve(self, d, x0=None, want_x=False):
        """(C column of load case d, solve statistics, converged[, solution])."""
        from chfem_amd.lib import SolveStats
        if not self.prepared:
            assert self.L.hmgPrepare() == 1
            self.prepared = True
        col = np.zeros(self.dim)
        ne, nd, _, _ = self.sizes()
        x = np.zeros(nd) if want_x else None
        x0p = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).ctypes.data_as(C.c_void_p)
        ok = self.L.hmgSolveLoadCase(d, x0p, None if x is None else x.ctypes.data_as(C.c_void_p),
                                     col.ctypes.data_as(C.c_void_p))
        st = SolveStats()
        self.L.cudapcgGetSolveStats(C.byref(st))
        return (col, st, ok == 1, x) if want_x else (col, st, ok == 1)

    def close(self):
        import os
        import shutil
        if self.prepared:
            self.L.hmgRelease()
            self.prepared = False
        if self.open:
            self.L.hmgEnd()
            self.open = False
            self.L.hmgSetPrecisionFlag(0)
            self.L.hmgSetParallelStrategyFlag(0)
            self.L.hmgSetSolverFlag(0)
        if os.path.isdir(self._dir):
            shutil.rmtree(self._dir)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
