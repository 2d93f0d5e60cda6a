"""The memory-lean solvers (CG3, MINRES3 with x, z and d; CG2, MINRES2 with
z and d and the solution reduced on the fly; chfem/src/femhmg/cudapcg/
solvers/{pcg3,pminres3,pcg2,pminres2,xreduce}.cu) and the INF and error
stopping rules (pcg.cu:130-146, pminres.cu:188-200) on the GPU, through the
C ABI, against the CPU FP64 oracle.

CG3 and MINRES3 are CG and MINRES with A d applied again instead of stored,
so their iterates are the oracle's CG / MINRES iterates: solutions within
1e-7 relative at tol 1e-10 and iteration counts within 2.  The stopping rules
are compared with the oracle's restatement (orc_solve_rule, itself pinned
against the true residual and the iterate change in tests/test_oracle.py):
the same iteration count within 1 and the rule's metric within tol.  The
two-vector solvers' returned values are the velocity sums of the oracle's
solution times the xreduce scale (FULL, ONLYDIR, NONE give the same total).
"""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from chfem_amd import synthetic
from tests import helpers as H

pytestmark = pytest.mark.gpu

CG, MINRES, CG3, MINRES3, CG2, MINRES2 = range(6)
L2, INF, ERR = 0, 1, 2


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def three_phase(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.choice(np.array([0, 100, 255], dtype=np.uint8), size=shape, p=[0.5, 0.3, 0.2])


@pytest.fixture(scope="module")
def thermal():
    return H.model(H.THERMAL, three_phase((7, 9, 11), 41), [1.0, 10.0, 3.5], colors=(0, 100, 255))


@pytest.fixture(scope="module")
def elastic():
    return H.model(H.ELASTIC, three_phase((6, 10, 8), 42), [1.0, 0.3, 10.0, 0.2, 4.0, 0.25],
                   colors=(0, 100, 255))


@pytest.fixture(scope="module")
def porous():
    img, _ = synthetic.porous_spheres((12, 20, 16), rmin=2, rmax=4, porosity=0.45, seed=4)
    return H.model(H.FLUID, img, [0.0], colors=(0, 255))


def gpu_solve(m, b, solver, tol, strategy=0, precond=1, resnorm=L2, maxit=20000, x0=None):
    s = H.GpuSolver(m, solver=solver, strategy=strategy, precond=precond, tol=tol, maxit=maxit)
    try:
        if resnorm != L2:
            assert s.L.cudapcgSetResNorm(resnorm) == 1
        x, ok, st = s.solve(b, x0=x0)
        return x, ok, st
    finally:
        s.close()


@pytest.mark.parametrize("precond", [1, 0], ids=["jacobi", "none"])
@pytest.mark.parametrize("strategy", [0, 1], ids=["nbn", "ebe"])
@pytest.mark.parametrize("solver", [CG3, MINRES3], ids=["cg3", "minres3"])
@pytest.mark.parametrize("which", ["thermal", "elastic"])
def test_three_vector_solvers_match_oracle(which, solver, strategy, precond, thermal, elastic):
    m = thermal if which == "thermal" else elastic
    for d in (0, m.cdim - 1):
        b = m.rhs(d)
        print(f"{which} solver {solver} strategy {strategy} precond {precond} load {d}", flush=True)
        x, ok, st = gpu_solve(m, b, solver, 1e-10, strategy=strategy, precond=precond)
        xr, okr, itr, _ = m.solve(d, solver=solver - 2, precond=precond, strategy=strategy, tol=1e-10)
        assert ok and okr
        assert rel(x, xr) < 1e-7
        assert abs(int(st.iterations) - itr) <= 2


def test_three_vector_generic_operator(elastic, monkeypatch):
    """The dense node-by-node operator has no lean epilogue: the solver runs
    the lean forms over one work vector."""
    monkeypatch.setenv("CHFEM_OPERATOR", "generic")
    m = elastic
    b = m.rhs(2)
    for solver in (CG3, MINRES3):
        x, ok, st = gpu_solve(m, b, solver, 1e-10)
        xr, okr, itr, _ = m.solve(2, solver=solver - 2, tol=1e-10)
        assert ok and okr and rel(x, xr) < 1e-7
        assert abs(int(st.iterations) - itr) <= 2


def test_three_vector_initial_guess(thermal):
    m = thermal
    b = m.rhs(1)
    xr, _, itr, _ = m.solve(1, tol=1e-10)
    # a guess halfway: fewer iterations, same solution
    x0 = 0.5 * xr
    for solver in (CG3, MINRES3):
        x, ok, st = gpu_solve(m, b, solver, 1e-10, x0=x0)
        assert ok and rel(x, xr) < 1e-7
        # the exact solution as guess: accepted without iterating
        x, ok, st = gpu_solve(m, b, solver, 1e-6, x0=xr)
        assert ok and st.iterations == 0 and rel(x, xr) < 1e-12


@pytest.mark.parametrize("solver", [CG3, MINRES3], ids=["cg3", "minres3"])
def test_three_vector_stokes(porous, solver):
    m = porous
    b = m.rhs(0)
    x, ok, st = gpu_solve(m, b, solver, 1e-9)
    xr, okr, itr, _ = m.solve(0, solver=solver - 2, tol=1e-9)
    assert ok and okr
    assert rel(m.update_c(0, x), m.update_c(0, xr)) < 1e-6
    assert abs(int(st.iterations) - itr) <= max(3, itr // 50)


@pytest.mark.parametrize("mode", [2, 1, 0], ids=["full", "onlydir", "none"])
@pytest.mark.parametrize("solver", [CG2, MINRES2], ids=["cg2", "minres2"])
def test_two_vector_reduction_modes(porous, solver, mode):
    """The reduced velocity of CG2 / MINRES2 in the three xreduce modes:
    FULL and NONE return every component's sum, ONLYDIR only the load
    case's, each times the xreduce scale."""
    m = porous
    d = 1
    xr, okr, _, _ = m.solve(d, solver=solver - 4, tol=1e-11, maxit=20000)
    assert okr
    sums = np.array([xr[a:3 * m.nvel:3].sum() for a in range(3)])
    scale = 1.0 / m.N
    s = H.GpuSolver(m, solver=solver, tol=1e-10)
    try:
        assert s.L.cudapcgSetXReduceMode(mode) == 1
        assert s.L.cudapcgSetXReduceShift(d) == 1
        assert s.L.cudapcgSetXReduceScale(scale) == 1
        assert s.L.cudapcgSetRHS(s.ptr(m.rhs(d))) == 1
        out = np.full(m.ndof, np.nan)
        assert s.L.cudapcgSolve(s.ptr(out)) == 1
    finally:
        s.close()
    want = scale * sums
    if mode == 1:
        want = np.where(np.arange(3) == d, want, 0.0)
    print(f"solver {solver} mode {mode}: {out[:3]} vs {want}")
    assert rel(out[:3], want) < 1e-7


@pytest.mark.parametrize("solver", [CG2, MINRES2], ids=["cg2", "minres2"])
def test_two_vector_permeability_and_metrics_file(solver):
    """Through the model layer: the permeability column of CG2 / MINRES2
    equals MINRES's, and each solve writes <model>_xreduce_metrics_<n>.bin
    with the iteration history of the scaled reduced velocity (iterations + 1
    values per component)."""
    img, _ = synthetic.porous_spheres((12, 14, 16), rmin=2, rmax=4, porosity=0.45, seed=7)
    with H.Resident(img, H.FLUID, {0: 0.0, 255: 0.0}, tol=1e-9, solver=MINRES) as r:
        ref = np.stack([r.solve(d)[0] for d in range(3)], axis=1)
    with H.Resident(img, H.FLUID, {0: 0.0, 255: 0.0}, tol=1e-9, solver=solver) as r:
        cols, its = [], []
        for d in range(3):
            col, st, ok = r.solve(d)
            assert ok
            cols.append(col)
            its.append(st.iterations)
        files = sorted(glob.glob(os.path.join(r._dir, "*_xreduce_metrics_*.bin")))
        assert len(files) == 3
        for f, it, col, d in zip(files, its, cols, range(3)):
            h = np.fromfile(f, dtype=np.float64).reshape(3, it + 1)
            assert np.all(h[:, 0] == 0.0)
            np.testing.assert_allclose(h[:, -1], col, rtol=1e-12)
    K = np.stack(cols, axis=1)
    print(f"solver {solver}: K {K.tolist()} vs MINRES {ref.tolist()}")
    assert rel(K, ref) < 1e-6


@pytest.mark.parametrize("rule", [INF, ERR], ids=["inf", "error"])
@pytest.mark.parametrize("solver", [CG, MINRES, CG3, MINRES3], ids=["cg", "minres", "cg3", "minres3"])
def test_stopping_rules_match_oracle(elastic, solver, rule):
    m = elastic
    tol = 1e-5 if rule == INF else 1e-4
    for d in (0, 5):
        b = m.rhs(d)
        x, ok, st = gpu_solve(m, b, solver, tol, resnorm=rule)
        xr, okr, itr, resr = m.solve_rule(d, rule, solver=solver % 2, tol=tol, maxit=20000)
        print(f"solver {solver} rule {rule} load {d}: {st.iterations} iterations (oracle {itr}), "
              f"metric {st.residual:.3e} (oracle {resr:.3e})", flush=True)
        assert ok == okr
        assert abs(int(st.iterations) - itr) <= 1
        if rule == INF:
            # the metric is the true residual's: check it through the oracle's operator
            true = np.abs(b - m.apply(x)).max() / np.abs(b).max()
            assert true <= tol * 1.001
        assert rel(x, xr) < 50 * tol


def test_stopping_rule_two_vector(porous):
    """MINRES2 with the INF rule: the oracle's MINRES INF count."""
    m = porous
    s = H.GpuSolver(m, solver=MINRES2, tol=1e-6)
    try:
        assert s.L.cudapcgSetResNorm(INF) == 1
        s.L.cudapcgSetXReduceScale(1.0)
        assert s.L.cudapcgSetRHS(s.ptr(m.rhs(2))) == 1
        out = np.zeros(m.ndof)
        s.L.cudapcgSolve(s.ptr(out))
        it = s.L.cudapcgGetNumIterations()
    finally:
        s.close()
    xr, okr, itr, _ = m.solve_rule(2, INF, solver=1, tol=1e-6, maxit=20000)
    assert okr and abs(int(it) - itr) <= 1
    sums = np.array([xr[a:3 * m.nvel:3].sum() for a in range(3)])
    assert rel(out[:3], sums) < 1e-4


def test_lean_solvers_hold_fewer_vectors(elastic):
    """Device memory the solver allocates: CG holds x, r, d, q, s, two CG
    loop buffers and the stored diagonal (FP64) before its direction ring;
    CG3 x, r and d; CG2 r and d (here through the raw ABI, elastic)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]

    def free_bytes():
        f, t = C.c_size_t(), C.c_size_t()
        assert hip.hipMemGetInfo(C.byref(f), C.byref(t)) == 0
        return f.value

    m = H.model(H.ELASTIC, three_phase((96, 96, 96), 43), [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    vec = m.ndof * 8
    used = {}
    for solver in (CG, CG3, CG2):
        s = H.GpuSolver(m, solver=solver, tol=1e-6)
        try:
            before = free_bytes()
            assert s.L.cudapcgAllocateArrays() == 1
            used[solver] = (before - free_bytes()) / vec
        finally:
            s.L.cudapcgFreeArrays()
            s.close()
    print(f"solver vectors allocated (units of one FP64 vector, {vec / 2**20:.1f} MiB): {used}")
    assert used[CG3] <= 2.5 and used[CG2] <= 1.5
    assert used[CG] >= used[CG3] + 3.5


@pytest.mark.parametrize("ring", ["2", "3", "5", "auto"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_minres_direction_ring(porous, ring, precision, monkeypatch):
    """MINRES with the L2 rule keeps its directions in a ring and adds
    alpha_j d_j to the iterate once per window: CHFEM_RING=2 leaves it no
    ring (the update kernel updates x), 3 gives 2 slots (a window flush every
    2 iterations, the tightest reuse), 5 gives 4, and the default 8.  The
    FP64 iterations and solution are the oracle's MINRES ones; the mixed
    solve (FP32 sweeps that end on stagnation, flushing their last window)
    reaches the same permeability column."""
    if ring != "auto":
        monkeypatch.setenv("CHFEM_RING", ring)
    d = 2
    xr, okr, itr, _ = porous.solve(d, solver=MINRES, tol=1e-10, maxit=20000)
    assert okr
    s = H.GpuSolver(porous, solver=MINRES, precision=precision, tol=1e-10 if precision == 0 else 1e-9,
                    maxit=20000)
    try:
        x, ok, st = s.solve(porous.rhs(d))
    finally:
        s.close()
    assert ok
    if precision == 0:
        assert abs(int(st.iterations) - itr) <= 2
        assert rel(x, xr) < 1e-7
    assert rel(porous.update_c(d, x), porous.update_c(d, xr)) < 1e-7


def test_minres_then_cg_same_solver(elastic):
    """The direction buffers are allocated on first use up to the loop's
    width: MINRES takes 9 slots, a later CG solve on the same solver grows
    them to its own width.  Both solves against the oracle's."""
    d = 1
    b = elastic.rhs(d)
    xm, okm, _, _ = elastic.solve(d, solver=MINRES, tol=1e-10, maxit=20000)
    xc, okc, itc, _ = elastic.solve(d, solver=CG, tol=1e-10, maxit=20000)
    assert okm and okc
    s = H.GpuSolver(elastic, solver=MINRES, tol=1e-10, maxit=20000)
    try:
        x1, ok1, _ = s.solve(b)
        assert ok1 and rel(x1, xm) < 1e-7
        assert s.L.cudapcgSetSolver(CG) == 1
        x2, ok2, st2 = s.solve(b)
        assert ok2 and rel(x2, xc) < 1e-7
        assert abs(int(st2.iterations) - itc) <= 2
    finally:
        s.close()
