"""GPU parity of the HIP path against the CPU FP64 oracle, through the C ABI.

Tolerances: operator applies, Jacobi, RHS and the effective-tensor column are
the same sums in a different order, so FP64 agreement is to 1e-12 relative
(1e-5 for the FP32 operator).  Solves compare effective tensors; both sides
iterate to the same L2 criterion, so agreement is bounded by the solver
tolerance (pinned here at 1e-9 -> C within 1e-7 relative).
"""
import numpy as np
import pytest

from chfem_amd import synthetic
from tests import helpers as H

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def three_phase(shape, seed):
    rng = np.random.default_rng(seed)
    img = rng.choice(np.array([0, 100, 255], dtype=np.uint8), size=shape, p=[0.5, 0.3, 0.2])
    return img


@pytest.fixture(scope="module")
def ragged_thermal():
    img = three_phase((7, 9, 11), 11)
    return H.model(H.THERMAL, img, [1.0, 10.0, 3.5], colors=(0, 100, 255))


@pytest.fixture(scope="module")
def ragged_elastic():
    img = three_phase((6, 10, 8), 12)
    return H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))


@pytest.fixture(scope="module")
def porous():
    img, _ = synthetic.porous_spheres((12, 20, 16), rmin=2, rmax=4, porosity=0.45, seed=4)
    return H.model(H.FLUID, img, [0.0], colors=(0, 255))


@pytest.mark.parametrize("strategy", [0, 1])
@pytest.mark.parametrize("which", ["thermal", "elastic"])
def test_operator_apply(which, strategy, ragged_thermal, ragged_elastic):
    m = ragged_thermal if which == "thermal" else ragged_elastic
    s = H.GpuSolver(m, strategy=strategy)
    try:
        rng = np.random.default_rng(5)
        for _ in range(2):
            v = rng.standard_normal(m.ndof)
            q_gpu = s.apply(v)
            q_ref = m.apply(v, strategy=0)
            assert rel(q_gpu, q_ref) < 1e-12
        z = s.precond(v)
        assert rel(z, m.jacobi_inv() * v) < 1e-14
    finally:
        s.close()


@pytest.mark.parametrize("kernel", ["walsh", "generic"])
def test_operator_fluid(porous, kernel, monkeypatch):
    """Stokes operator (Walsh-basis tiles by default, the node-by-node
    kernel with CHFEM_OPERATOR=generic) and Jacobi against the oracle."""
    if kernel == "generic":
        monkeypatch.setenv("CHFEM_OPERATOR", "generic")
    m = porous
    assert m.nvel > 0 and m.nborder > 0
    s = H.GpuSolver(m)
    try:
        v = np.random.default_rng(6).standard_normal(m.ndof)
        assert rel(s.apply(v), m.apply(v)) < 1e-12
        assert rel(s.precond(v), m.jacobi_inv() * v) < 1e-14
    finally:
        s.close()


@pytest.mark.parametrize("kernel", ["default", "generic"])
@pytest.mark.parametrize("which", ["thermal", "elastic", "fluid"])
def test_operator_fp32(which, kernel, ragged_thermal, ragged_elastic, porous, monkeypatch):
    """The FP32 operators of the mixed-precision solver against the oracle
    (FP64) on an FP32-representable input: every entry within 1e-5 of the
    largest (FP32 rounding of sums over 8 elements)."""
    if kernel == "generic":
        monkeypatch.setenv("CHFEM_OPERATOR", "generic")
    m = {"thermal": ragged_thermal, "elastic": ragged_elastic, "fluid": porous}[which]
    s = H.GpuSolver(m)
    try:
        v = np.random.default_rng(9).standard_normal(m.ndof).astype(np.float32)
        q32 = s.apply32(v)
        ref = m.apply(v.astype(np.float64))
        assert np.abs(q32 - ref).max() < 1e-5 * np.abs(ref).max()
    finally:
        s.close()


@pytest.mark.parametrize("kernel", ["walsh", "generic"])
def test_fluid_mixed_precision(porous, kernel, monkeypatch):
    """Permeability load case with FP32 inner MINRES and FP64 refinement,
    against the oracle's FP64 solve."""
    if kernel == "generic":
        monkeypatch.setenv("CHFEM_OPERATOR", "generic")
    m = porous
    s = H.GpuSolver(m, solver=1, precision=1, tol=1e-8)
    try:
        x, ok, st = s.solve(m.rhs(1))
        xr, okr, _, _ = m.solve(1, solver=1, tol=1e-10)
        assert ok and okr, (st.iterations, st.sweeps, st.residual)
        assert rel(m.update_c(1, x), m.update_c(1, xr)) < 1e-6
    finally:
        s.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("shape", [(40, 70, 9), (17, 33, 65), (3, 5, 4), (9, 130, 14)])
def test_operator_fluid_tiles(shape, precision, monkeypatch):
    """Walsh-basis Stokes operator on ragged, multi-block and tiny grids with
    random pore spaces (several tiles per direction, layer chunks), against
    the oracle: FP64 to 1e-12, FP32 on an FP32-representable input to 1e-5
    of the largest entry."""
    monkeypatch.setenv("CHFEM_WALSH_TILE", "32x8x7")
    rng = np.random.default_rng(sum(shape))
    img = np.where(rng.random(shape) < 0.35, 255, 0).astype(np.uint8)
    # the 8 elements around node (1, 1, 1) are pore, so it is fully fluid
    img[:2, :2, :2] = 0
    m = H.model(H.FLUID, img, [0.0], colors=(0, 255))
    assert m.nvel > 0 and m.nborder > 0
    s = H.GpuSolver(m)
    try:
        v = rng.standard_normal(m.ndof)
        if precision == "fp64":
            assert rel(s.apply(v), m.apply(v)) < 1e-12
        else:
            v32 = v.astype(np.float32)
            ref = m.apply(v32.astype(np.float64))
            assert np.abs(s.apply32(v32) - ref).max() < 1e-5 * np.abs(ref).max()
    finally:
        s.close()


def test_pcg_matches_oracle_thermal(ragged_thermal):
    m = ragged_thermal
    s = H.GpuSolver(m, tol=1e-10)
    try:
        for d in range(3):
            b = m.rhs(d)
            x, ok, st = s.solve(b)
            xr, okr, itr, _ = m.solve(d, tol=1e-10)
            assert ok and okr
            assert rel(x, xr) < 1e-7
            # same algorithm, same stop rule: iteration counts agree to reduction-order noise
            assert abs(int(st.iterations) - itr) <= 2
    finally:
        s.close()


def test_pcg_matches_oracle_elastic_ebe(ragged_elastic):
    m = ragged_elastic
    s = H.GpuSolver(m, strategy=1, tol=1e-10)
    try:
        b = m.rhs(3)
        x, ok, st = s.solve(b)
        xr, okr, itr, _ = m.solve(3, tol=1e-10)
        assert ok and okr and rel(x, xr) < 1e-7
    finally:
        s.close()


@pytest.mark.parametrize("solver", [0, 1])
def test_fluid_solvers_match_oracle(porous, solver):
    m = porous
    s = H.GpuSolver(m, solver=solver, tol=1e-9)
    try:
        b = m.rhs(0)
        x, ok, st = s.solve(b)
        xr, okr, itr, _ = m.solve(0, solver=solver, tol=1e-9)
        assert ok and okr
        C = m.update_c(0, x)
        Cr = m.update_c(0, xr)
        assert rel(C, Cr) < 1e-6
        assert abs(int(st.iterations) - itr) <= max(3, itr // 50)
    finally:
        s.close()


def test_initial_guess_path(ragged_thermal):
    m = ragged_thermal
    s = H.GpuSolver(m, tol=1e-8)
    try:
        b = m.rhs(1)
        xr, _, itr, _ = m.solve(1, tol=1e-10)
        x, ok, st = s.solve(b, x0=xr)
        assert ok and st.iterations <= 1
        assert rel(x, xr) < 1e-8
    finally:
        s.close()


def test_mixed_precision_elastic(ragged_elastic):
    m = ragged_elastic
    s = H.GpuSolver(m, precision=1, tol=1e-8)
    try:
        b = m.rhs(0)
        x, ok, st = s.solve(b)
        xr, okr, _, _ = m.solve(0, tol=1e-10)
        assert ok
        assert rel(m.update_c(0, x), m.update_c(0, xr)) < 1e-7
        assert st.sweeps >= 1
    finally:
        s.close()


@pytest.mark.parametrize("tile", ["32x8x16", "16x16x4", "64x4x7", "32x16x3", "64x8x5", "65x7x5", "65x15x3"])
@pytest.mark.parametrize("shape", [(40, 70, 9), (3, 1, 2), (17, 33, 65)])
@pytest.mark.parametrize("which", ["elastic", "thermal"])
def test_walsh_operator_tiles(which, tile, shape, monkeypatch):
    """Walsh-basis operators (isotropic elastic, conduction) on ragged, tiny
    and multi-chunk grids for every tile shape, NBN and EBE (atomic-form
    strategy, which also routes to the Walsh kernel), against the oracle's
    dense local matrices."""
    monkeypatch.setenv("CHFEM_WALSH_TILE", tile)
    monkeypatch.setenv("CHFEM_EBE", "atomic")
    img = three_phase(shape, 21)
    if which == "elastic":
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    else:
        m = H.model(H.THERMAL, img, [1.0, 10.0, 3.5], colors=(0, 100, 255))
    for strategy in (0, 1):
        s = H.GpuSolver(m, strategy=strategy)
        try:
            v = np.random.default_rng(7).standard_normal(m.ndof)
            assert rel(s.apply(v), m.apply(v)) < 1e-12
        finally:
            s.close()


@pytest.mark.parametrize("which", ["elastic", "thermal"])
def test_generic_operator_path(which, ragged_elastic, ragged_thermal, monkeypatch):
    """The dense per-material kernels (24x24 elastic, 8x8 conduction; used for
    local matrices that are not isotropic / unit conduction matrices)."""
    monkeypatch.setenv("CHFEM_OPERATOR", "generic")
    m = ragged_elastic if which == "elastic" else ragged_thermal
    s = H.GpuSolver(m)
    try:
        v = np.random.default_rng(8).standard_normal(m.ndof)
        assert rel(s.apply(v), m.apply(v)) < 1e-12
    finally:
        s.close()


@pytest.mark.parametrize("precision", [0, 1])
def test_walsh_solve_multiblock(precision, monkeypatch):
    """Solve with the Walsh operator on a grid of several blocks along each of
    rows, columns and layers, so its fused d.Ad reduction runs on a 3-D grid."""
    monkeypatch.setenv("CHFEM_WALSH_TILE", "16x16x4")
    img = three_phase((14, 40, 33), 31)
    m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    s = H.GpuSolver(m, precision=precision, tol=1e-9)
    try:
        b = m.rhs(4)
        x, ok, st = s.solve(b)
        xr, okr, itr, _ = m.solve(4, tol=1e-9)
        assert ok and okr
        assert rel(m.update_c(4, x), m.update_c(4, xr)) < 1e-7
        if precision == 0:
            assert abs(int(st.iterations) - itr) <= 2
    finally:
        s.close()


@pytest.mark.parametrize("vector", ["keys", "stored"])
@pytest.mark.parametrize("which", ["thermal", "elastic"])
def test_pcg_vector_kernels(which, vector, monkeypatch):
    """Both PCG vector-kernel forms (Jacobi diagonal rebuilt from the keys, or
    a stored dinv vector) on a grid whose node count is not a multiple of 4
    (tail nodes), in FP64 and mixed precision, against the oracle."""
    if vector == "stored":
        monkeypatch.setenv("CHFEM_VECTOR", "stored")
    img = three_phase((5, 7, 9), 41)
    if which == "thermal":
        m = H.model(H.THERMAL, img, [1.0, 10.0, 3.5], colors=(0, 100, 255))
    else:
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    assert m.N % 4 == 3
    d = m.cdim - 1
    xr, okr, itr, _ = m.solve(d, tol=1e-10)
    for precision in (0, 1):
        s = H.GpuSolver(m, precision=precision, tol=1e-10 if precision == 0 else 1e-9)
        try:
            x, ok, st = s.solve(m.rhs(d))
            assert ok and okr
            assert rel(m.update_c(d, x), m.update_c(d, xr)) < 1e-7
            if precision == 0:
                assert rel(x, xr) < 1e-7
                assert abs(int(st.iterations) - itr) <= 2
        finally:
            s.close()


def test_mfma_f32_f64_layout():
    """Lane layout of the 16x16x4 FP32/FP64 matrix-core products the EBE
    operator is built on."""
    from chfem_amd.lib import lib, ptr
    rng = np.random.default_rng(1)
    A, B = rng.standard_normal((16, 4)), rng.standard_normal((4, 16))
    Af, Bf = A.astype(np.float32), B.astype(np.float32)
    D = np.zeros((16, 16), np.float32)
    D2 = np.zeros((16, 16))
    assert lib().chfemMfmaCheckF(ptr(Af), ptr(Bf), ptr(D), ptr(A), ptr(B), ptr(D2)) == 1
    np.testing.assert_allclose(D, Af.astype(np.float64) @ Bf.astype(np.float64), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(D2, A @ B, rtol=1e-13, atol=1e-13)


def test_mfma_4x4_layout():
    """Lane layout of the FP32 4x4x1 16-block matrix-core product the FP32
    EBE operator is built on: lane l = 4 b + j supplies row j of block b's A
    and column j of its B, and holds column j of D_b = A_b B_b (rows 0..3
    in its 4 result items)."""
    from chfem_amd.lib import lib, ptr
    rng = np.random.default_rng(2)
    A = rng.standard_normal(64).astype(np.float32)
    B = rng.standard_normal(64).astype(np.float32)
    D = np.zeros(256, np.float32)
    assert lib().chfemMfmaCheck4(ptr(A), ptr(B), ptr(D)) == 1
    Ab, Bb = A.reshape(16, 4).astype(np.float64), B.reshape(16, 4).astype(np.float64)
    Db = np.einsum("bi,bj->bij", Ab, Bb)  # D_b[i][j] = A_b[i] B_b[j]
    got = D.reshape(16, 4, 4)  # [block][column j][row i]
    np.testing.assert_allclose(got.transpose(0, 2, 1), Db, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("kernel", ["mfma", "mfma-32x8", "atomic"])
@pytest.mark.parametrize("which", ["thermal", "elastic"])
@pytest.mark.parametrize("nphase", [1, 4])
def test_ebe_kernels(kernel, which, nphase, monkeypatch):
    """EBE operator on the matrix cores (16x16 element tiles by default, and
    32x8) and the reference-style element kernel with atomics, with 1 and 4
    materials on a grid of several blocks in every direction, FP64 and FP32
    against the oracle."""
    if kernel == "atomic":
        monkeypatch.setenv("CHFEM_EBE", "atomic")
    if kernel == "mfma-32x8":
        monkeypatch.setenv("CHFEM_EBE_TILE", "32x8")
    rng = np.random.default_rng(51 + nphase)
    colors = (0, 60, 120, 255)[:nphase]
    img = rng.choice(np.array(colors, dtype=np.uint8), size=(20, 40, 35))
    if which == "thermal":
        m = H.model(H.THERMAL, img, [1.0, 10.0, 3.5, 0.5][:nphase], colors=colors)
    else:
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25, 2.0, 0.35][:2 * nphase], colors=colors)
    v = rng.standard_normal(m.ndof)
    q_ref = m.apply(v, strategy=1)
    v32 = v.astype(np.float32)
    q_ref32 = m.apply(v32.astype(np.float64), strategy=1)
    for precision in (0, 1):
        s = H.GpuSolver(m, strategy=1, precision=precision, tol=1e-9)
        try:
            if precision == 0:
                assert rel(s.apply(v), q_ref) < 1e-12
            else:
                # FP32 operator on an FP32-representable input: FP32 rounding of
                # sums over 8 elements
                q32 = s.apply32(v32)
                assert np.abs(q32 - q_ref32).max() < 1e-5 * np.abs(q_ref32).max()
                x, ok, st = s.solve(m.rhs(0))
                xr, okr, _, _ = m.solve(0, tol=1e-10, strategy=1)
                assert ok and okr
                assert rel(m.update_c(0, x), m.update_c(0, xr)) < 1e-7
        finally:
            s.close()


@pytest.mark.parametrize("which", ["thermal", "elastic"])
def test_ebe_many_materials(which):
    """More than 4 materials (7 thermal, 5 elastic): the EBE operator runs
    the element kernel with atomics, since the matrix-core kernel keeps at
    most 4 local matrices in LDS.  FP64 and FP32 operators and the mixed
    solve's C column against the oracle."""
    rng = np.random.default_rng(77)
    if which == "thermal":
        colors = (0, 40, 80, 120, 160, 200, 255)
        props = [1.0, 10.0, 3.5, 0.5, 2.0, 7.0, 4.5]
        an = H.THERMAL
    else:
        colors = (0, 60, 120, 180, 255)
        props = [1.0, 0.3, 10.0, 0.2, 4.0, 0.25, 2.0, 0.35, 6.0, 0.28]
        an = H.ELASTIC
    img = rng.choice(np.array(colors, dtype=np.uint8), size=(12, 18, 15))
    m = H.model(an, img, props, colors=colors)
    v = rng.standard_normal(m.ndof)
    q_ref = m.apply(v, strategy=1)
    v32 = v.astype(np.float32)
    q_ref32 = m.apply(v32.astype(np.float64), strategy=1)
    xr, okr, _, _ = m.solve(1, tol=1e-10, strategy=1)
    assert okr
    for precision in (0, 1):
        s = H.GpuSolver(m, strategy=1, precision=precision, tol=1e-9)
        try:
            if precision == 0:
                assert rel(s.apply(v), q_ref) < 1e-12
            else:
                q32 = s.apply32(v32)
                assert np.abs(q32 - q_ref32).max() < 1e-5 * np.abs(q_ref32).max()
            x, ok, st = s.solve(m.rhs(1))
            assert ok
            assert rel(m.update_c(1, x), m.update_c(1, xr)) < 1e-7
        finally:
            s.close()


@pytest.mark.parametrize("which", ["elastic", "thermal"])
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("fuse", ["2", "1", "0"])
def test_fused_direction_update(fuse, precision, which, monkeypatch):
    """The CG loop forms on the Walsh operators: single-kernel iterations
    with the z update (and the r.z recurrence) inside the operator
    (CHFEM_FUSE_DIR=2, the default), the direction and x updates inside the
    operator with a z update kernel (1), and separate kernels (0), on a grid of several
    blocks in every direction, from the zero guess and from a non-zero guess,
    in FP64 and mixed precision (FP64 residual replacements keep the
    direction), against the oracle."""
    monkeypatch.setenv("CHFEM_FUSE_DIR", fuse)
    monkeypatch.setenv("CHFEM_WALSH_TILE", "16x16x4")
    img = three_phase((21, 34, 19), 51)
    if which == "elastic":
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    else:
        m = H.model(H.THERMAL, img, [1.0, 10.0, 3.5], colors=(0, 100, 255))
    xr, okr, itr, _ = m.solve(2, tol=1e-10)
    assert okr
    Cr = m.update_c(2, xr)
    s = H.GpuSolver(m, precision=precision, tol=1e-10 if precision == 0 else 1e-9)
    try:
        x, ok, st = s.solve(m.rhs(2))
        assert ok
        assert rel(m.update_c(2, x), Cr) < 1e-7
        if precision == 0:
            assert rel(x, xr) < 1e-7
            assert abs(int(st.iterations) - itr) <= 2
        else:
            assert st.sweeps >= 2
        x1, ok1, st1 = s.solve(m.rhs(2), x0=0.5 * xr)
        assert ok1 and st1.x0_used == 1
        assert rel(m.update_c(2, x1), Cr) < 1e-7
        assert st1.iterations < st.iterations + 2
    finally:
        s.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("fuse", ["2", "1"])
@pytest.mark.parametrize("ring", ["2", "3", "4", "5", "16"])
def test_direction_ring_widths(ring, fuse, precision, monkeypatch):
    """The fused CG loops keep their directions in a ring of CHFEM_RING
    buffers and add sum alpha_j d_j to the iterate once per ring window and
    at every loop or FP64-replacement end.  Windows that end inside a batch
    of queued iterations, at a replacement, or after the loop has converged
    (queued iterations that did not run) all give the oracle's solution
    (widths 2 and 5 end windows inside the 16-iteration batches).  Width 3
    is the smallest at which the single-kernel iteration recovers z from the
    two previous directions: iteration j writes slot j mod 3 while it reads
    slots j - 1 and j - 2, and every flush window is the three slots in use."""
    monkeypatch.setenv("CHFEM_FUSE_DIR", fuse)
    monkeypatch.setenv("CHFEM_RING", ring)
    monkeypatch.setenv("CHFEM_WALSH_TILE", "16x16x4")
    img = three_phase((21, 34, 19), 53)
    m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    xr, okr, itr, _ = m.solve(1, tol=1e-10)
    assert okr
    Cr = m.update_c(1, xr)
    s = H.GpuSolver(m, precision=precision, tol=1e-10 if precision == 0 else 1e-9)
    try:
        x, ok, st = s.solve(m.rhs(1))
        assert ok
        assert rel(m.update_c(1, x), Cr) < 1e-7
        if precision == 0:
            assert rel(x, xr) < 1e-7
            assert abs(int(st.iterations) - itr) <= 2
    finally:
        s.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("fuse", ["2", "1"])
@pytest.mark.parametrize("tile", ["65x7x5", "65x15x3"])
@pytest.mark.parametrize("which", ["elastic", "thermal"])
def test_halo_tiles_fused_solve(which, tile, fuse, precision, monkeypatch):
    """Tiles of 64 owned rows (65 element rows: the row above each tile is
    evaluated by an extra wavefront) in the fused CG loops: two tiles along
    the rows (70 = 64 + 6), ragged columns and layers; the solution's
    effective column against the oracle's, and the FP64 iterate itself."""
    monkeypatch.setenv("CHFEM_FUSE_DIR", fuse)
    monkeypatch.setenv("CHFEM_WALSH_TILE", tile)
    img = three_phase((11, 70, 23), 61)
    if which == "elastic":
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    else:
        m = H.model(H.THERMAL, img, [1.0, 10.0, 3.5], colors=(0, 100, 255))
    xr, okr, itr, _ = m.solve(0, tol=1e-10)
    assert okr
    s = H.GpuSolver(m, precision=precision, tol=1e-10 if precision == 0 else 1e-9)
    try:
        x, ok, st = s.solve(m.rhs(0))
        assert ok
        assert rel(m.update_c(0, x), m.update_c(0, xr)) < 1e-7
        if precision == 0:
            assert rel(x, xr) < 1e-7
            assert abs(int(st.iterations) - itr) <= 2
    finally:
        s.close()


@pytest.mark.parametrize("which", ["thermal", "elastic"])
@pytest.mark.parametrize("shape", [(8, 10, 12), (7, 9, 11)])
def test_coarse_to_fine_interpolation(which, shape):
    """cudapcgSetX0(x0, mustInterpolate=1) (k_project2 / k_interp) against
    the oracle's restatement of image.cu:19-121 (orc_interpl2).  The coarse
    vector is the oracle's solution on the 2x coarser mesh (the recursive
    guess's input), so the interpolated guess beats the zero guess and a
    solve at a loose tolerance returns it unchanged."""
    from oracle import oracle as O
    rng = np.random.default_rng(sum(shape))
    img = three_phase(shape, 71)
    nz, ny, nx = shape
    if which == "thermal":
        m = H.model(H.THERMAL, img, [1.0, 10.0, 3.5], colors=(0, 100, 255))
    else:
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], colors=(0, 100, 255))
    nd = 1 if which == "thermal" else 3
    ce = np.zeros((nx // 2) * (ny // 2) * (nz // 2), np.uint16)
    O.lib().orc_coarse_elem_map(O._p(m.elem), nx, ny, nz, O._p(ce))
    mc = O.Model(m.analysis, (nx // 2, ny // 2, nz // 2), ce, m.props[:2 * m.nmat] if nd == 3 else m.props[:m.nmat])
    xc, okc, _, _ = mc.solve(0, tol=1e-10)
    assert okc
    fine = np.zeros(m.ndof)
    O.lib().orc_interpl2(O._p(xc), ny, nx, nz, nd, O._p(fine))
    s = H.GpuSolver(m, tol=1e3)
    try:
        assert s.L.cudapcgSetRHS(s.ptr(m.rhs(0))) == 1
        assert s.L.cudapcgSetX0(s.ptr(xc), 1) == 1
        x = np.zeros(m.ndof)
        s.L.cudapcgSolve(s.ptr(x))
        from chfem_amd.lib import SolveStats
        import ctypes
        st = SolveStats()
        s.L.cudapcgGetSolveStats(ctypes.byref(st))
        assert st.x0_used == 1 and st.iterations == 0
        assert rel(x, fine) < 1e-14
    finally:
        s.close()


@pytest.mark.parametrize("solver", [4, 5], ids=["cg2", "minres2"])
def test_reduced_velocity_solvers(porous, solver):
    """CG2 / MINRES2 (two-vector solvers for permeability) return the
    effective velocity of the load case instead of the solution vector: the
    sum of each velocity component over the velocity DOFs times
    voxel^2 / nelem (xreduce.cu:4-26, femhmg.c:922-923,1087-1093), which is
    the column the effective-tensor integration gives from the full
    solution."""
    m = porous
    xr, okr, _, _ = m.solve(2, solver=solver - 4, tol=1e-10)
    assert okr
    Cref = m.update_c(2, xr).reshape(3, 3)[:, 2]
    s = H.GpuSolver(m, solver=solver, tol=1e-10)
    try:
        s.L.cudapcgSetXReduceScale(1.0 / m.N)
        assert s.L.cudapcgSetRHS(s.ptr(m.rhs(2))) == 1
        out = np.zeros(m.ndof)
        assert s.L.cudapcgSolve(s.ptr(out)) == 1
        assert rel(out[:3], Cref) < 1e-7
    finally:
        s.close()
