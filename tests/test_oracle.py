"""Pins the CPU FP64 oracle (oracle/chfem_oracle.c) before it is trusted as the
parity reference:

* the reference's own golden results (chfem test/test_chfem.py:6-22,
  test/test_chfem.c:340-351: 1e-4 relative to max |C|);
* brute-force global assembly in NumPy (tests/brute.py) for the keys, fluid
  DOF numbering, operator, Jacobi diagonal, right-hand sides and effective
  tensor columns;
* closed forms: energies of linear fields in the local matrices, homogeneous
  images (C = material tensor), laminates (exact Voigt/Reuss-type averages),
  reflections/axis swaps of the image (tensor transformation of C), CG vs
  MINRES vs a dense least-squares solve, coarse-to-fine interpolation of
  linear fields.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests import brute as B
from tests.conftest import GOLDEN

THERMAL, ELASTIC, FLUID = O.THERMAL, O.ELASTIC, O.FLUID


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def elem_of(img, colors):
    """Element map from a (z, y, x) image: element r + c*ny + l*nx*ny takes the
    material index of voxel (z=l, y=r, x=c) (chfem/src/femhmg/femhmg.c:1663-1669)."""
    lut = np.zeros(256, dtype=np.uint16)
    for k, c in enumerate(colors):
        lut[c] = k
    return np.ascontiguousarray(lut[img].transpose(0, 2, 1)).reshape(-1)


def make(analysis, img, props, colors=(0, 255)):
    nz, ny, nx = img.shape
    return O.Model(analysis, (nx, ny, nz), elem_of(img, colors), props)


def three_phase(shape, seed, p=(0.5, 0.3, 0.2)):
    rng = np.random.default_rng(seed)
    return rng.choice(np.array([0, 100, 255], dtype=np.uint8), size=shape, p=list(p))


def local(m):
    K, CB = m.local()
    nd = {THERMAL: 1, ELASTIC: 3, FLUID: 4}[m.analysis]
    if m.analysis == FLUID:
        return K.reshape(32, 32), CB
    ke = 8 * nd
    nrow = 3 if nd == 1 else 6
    return K.reshape(-1, ke, ke), CB.reshape(-1, nrow, ke)


# ---------------------------------------------------------------------------
# golden results held by the reference's tests
# ---------------------------------------------------------------------------
GOLDEN_CASES = [
    ("thermal_ggg40_100.nf", "ggg40_100.raw", (100, 100, 100), 0, "thermal_ggg40_100_result.bin", 3),
    ("elastic_ggg40_100.nf", "ggg40_100.raw", (100, 100, 100), 1, "elastic_ggg40_100_result.bin", 6),
    ("fluid_fibers_100x100x10.nf", "fibers_100x100x10.raw", (10, 100, 100), 2,
     "fluid_fibers_100x100x10_result.bin", 3),
]


@pytest.mark.slow
@pytest.mark.parametrize("case", GOLDEN_CASES, ids=["thermal", "elastic", "fluid"])
def test_oracle_golden(case):
    """Single-column runs of the reference's three 3D tests (thermal X,
    elastic Y, permeability Z with refinement 2), compared as the reference's
    C harness does: max |C - C_ref| / max |C_ref| < 1e-4 over the column."""
    nf, raw, shape, d, res, dim = case
    img = np.fromfile(os.path.join(GOLDEN, raw), dtype=np.uint8).reshape(shape)
    m, info = O.model_from_nf(os.path.join(GOLDEN, nf), img)
    C, it = m.homogenize(directions=[d], tol=info["tol"], maxit=info["maxit"], solver=info.get("solver", 0))
    ref = np.fromfile(os.path.join(GOLDEN, res)).reshape(dim, dim)
    assert it[d][0]
    assert np.abs(C[:, d] - ref[:, d]).max() / np.abs(ref).max() < 1e-4


# ---------------------------------------------------------------------------
# maps
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("ref", [1, 2, 3])
def test_image_to_elem_map(ref):
    """File order (z, y, x) -> element order with each voxel replicated ref^3
    times (chfem/src/femhmg/femhmg.c:1644-1758)."""
    img = three_phase((3, 4, 5), 1)
    keys = np.zeros(256, dtype=np.uint16)
    keys[[0, 100, 255]] = [2, 0, 1]
    em, dims = O.image_to_elem_map(img, {0: 2, 100: 0, 255: 1}, ref)
    fine = img.repeat(ref, 0).repeat(ref, 1).repeat(ref, 2)
    assert dims == (5 * ref, 4 * ref, 3 * ref)
    np.testing.assert_array_equal(em, keys[fine.transpose(0, 2, 1)].reshape(-1))


@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 1, 1), (2, 1, 3), (5, 3, 2)])
def test_packed_keys(shape):
    img = three_phase(shape, 2)
    m = make(ELASTIC, img, [1, 0.3, 2, 0.2, 3, 0.25], (0, 100, 255))
    np.testing.assert_array_equal(m.keys(), B.packed_keys(m.elem, m.nx, m.ny, m.nz))


@pytest.mark.parametrize("shape,seed", [((4, 5, 6), 3), ((3, 3, 3), 4), ((2, 5, 4), 5)])
def test_fluid_numbering(shape, seed):
    rng = np.random.default_rng(seed)
    img = np.where(rng.random(shape) < 0.7, 0, 255).astype(np.uint8)
    m = make(FLUID, img, [0.0])
    fk, ids, d2n, bk = m.fluid_maps()
    fk_b, ids_b, d2n_b, nvel, nb = B.fluid_maps(m.elem, m.nx, m.ny, m.nz)
    assert (m.nvel, m.nborder) == (nvel, nb)
    assert m.ndof == 4 * nvel + nb
    np.testing.assert_array_equal(fk, fk_b)
    np.testing.assert_array_equal(ids, ids_b)
    np.testing.assert_array_equal(d2n, d2n_b)
    np.testing.assert_array_equal(bk, fk_b[d2n_b[nvel:]])


# ---------------------------------------------------------------------------
# local matrices: closed forms
# ---------------------------------------------------------------------------
def frame(stokes=False):
    """Node coordinates of the unit voxel: x = column offset, y = 1 - row
    offset (rows run top to bottom), z = 1 - layer offset for thermal and
    elastic, z = layer offset for Stokes."""
    z = B.DL if stokes else 1 - B.DL
    return np.stack([B.DC, 1 - B.DR, z], axis=1).astype(float)


def test_thermal_local_closed_form():
    m = make(THERMAL, np.zeros((2, 2, 2), np.uint8), [2.5])
    K, CB = local(m)
    K = K[0] / 2.5
    X = frame()
    d2 = ((X[:, None] - X[None]) ** 2).sum(-1)
    # unit-cube Q1 Laplacian: 1/3 on the diagonal, 0 along edges, -1/12 across
    # face and body diagonals
    expect = np.select([d2 == 0, d2 == 1, d2 >= 2], [1 / 3, 0.0, -1 / 12])
    np.testing.assert_allclose(K, expect, atol=1e-15)
    for g in np.eye(3):
        T = X @ g
        assert abs(T @ K @ T - 1.0) < 1e-14
        # CB is k times the centre gradient
        np.testing.assert_allclose(CB[0] @ T, 2.5 * g, atol=1e-14)


def voigt_strain(u_grad):
    e = 0.5 * (u_grad + u_grad.T)
    return np.array([e[0, 0], e[1, 1], e[2, 2], 2 * e[1, 2], 2 * e[0, 2], 2 * e[0, 1]])


@pytest.mark.parametrize("E,nu", [(1.0, 0.3), (10.0, 0.2), (3.0, 0.0)])
def test_elastic_local_closed_form(E, nu):
    m = make(ELASTIC, np.zeros((2, 2, 2), np.uint8), [E, nu])
    K, CB = local(m)
    K, CB = K[0], CB[0]
    D = B.isotropic_C(E, nu)
    X = frame()
    rng = np.random.default_rng(0)
    np.testing.assert_allclose(K, K.T, atol=1e-14)
    for _ in range(8):
        G = rng.standard_normal((3, 3))  # u(x) = G x
        u = (X @ G.T).reshape(-1)
        eps = voigt_strain(G)
        assert abs(u @ K @ u - eps @ D @ eps) < 1e-12 * abs(eps @ D @ eps)
        np.testing.assert_allclose(CB @ u, D @ eps, atol=1e-12)
    # 6 rigid-body modes, 18 deformation modes
    w = np.linalg.eigvalsh(K)
    assert (np.abs(w) < 1e-12).sum() == 6 and (w > 1e-6).sum() == 18


def test_stokes_local_closed_form():
    m = make(FLUID, np.zeros((2, 2, 2), np.uint8), [0.0])
    K, CB = local(m)
    X = frame(stokes=True)
    Kt = make(THERMAL, np.zeros((2, 2, 2), np.uint8), [1.0]).local()[0].reshape(8, 8)
    rng = np.random.default_rng(1)
    for _ in range(6):
        G = rng.standard_normal((3, 3))
        u = (X @ G.T).reshape(-1)
        e = 0.5 * (G + G.T)
        # viscous block: 2 mu eps:eps with mu = 1
        assert abs(u @ K[:24, :24] @ u - 2 * (e * e).sum()) < 1e-12
        # coupling block: constant pressure tests the divergence
        assert abs(u @ K[:24, 24:] @ np.ones(8) - np.trace(G)) < 1e-13
    np.testing.assert_allclose(K[24:, :24], K[:24, 24:].T, atol=1e-15)
    # stabilisation: tau grad.grad with tau = 1/6 for the unit voxel
    np.testing.assert_allclose(K[24:, 24:], Kt / 6.0, atol=1e-15)
    np.testing.assert_allclose(CB[72:], 0.125)


def test_stokes_jacobi_constants():
    """Jacobi of the stabilised Stokes system: 1/(32/9) on velocity rows,
    1/(-4/9) on pore-pressure rows, -18/popcount(key) on border rows
    (chfem/src/femhmg/cudapcg/kernels/matrix_free.cu:663-672,912-925)."""
    rng = np.random.default_rng(7)
    img = np.where(rng.random((4, 5, 6)) < 0.7, 0, 255).astype(np.uint8)
    m = make(FLUID, img, [0.0])
    d = m.jacobi_inv()
    nv, nb = m.nvel, m.nborder
    np.testing.assert_allclose(d[:3 * nv], 0.28125, rtol=1e-14)
    np.testing.assert_allclose(d[3 * nv:4 * nv], -2.25, rtol=1e-14)
    _, _, _, bk = m.fluid_maps()
    pop = np.array([bin(int(k)).count("1") for k in bk])
    np.testing.assert_allclose(d[4 * nv:], -18.0 / pop, rtol=1e-14)


# ---------------------------------------------------------------------------
# operator, Jacobi, RHS and effective column against brute-force assembly
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("analysis", [THERMAL, ELASTIC])
@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 2, 3), (4, 1, 1)])
def test_operator_brute(analysis, shape):
    img = three_phase(shape, 8)
    props = [1.0, 10.0, 3.5] if analysis == THERMAL else [1.0, 0.3, 10.0, 0.2, 4.0, 0.25]
    m = make(analysis, img, props, (0, 100, 255))
    Kl, _ = local(m)
    nd = 1 if analysis == THERMAL else 3
    A = B.assemble(m.elem, m.nx, m.ny, m.nz, Kl, nd)
    v = np.random.default_rng(9).standard_normal(m.ndof)
    Av = A @ v
    assert rel(m.apply(v, strategy=0), Av) < 1e-13
    assert rel(m.apply(v, strategy=1), Av) < 1e-13
    assert abs(A - A.T).max() < 1e-13
    # Jacobi sums the local diagonals of the 8 elements around a node; with a
    # grid dimension of 1 the global diagonal also picks up couplings of a node
    # with its own periodic image, which the reference's diagonal leaves out
    conn = B.connectivity(m.nx, m.ny, m.nz)
    diag = np.zeros(m.ndof)
    for a in range(nd):
        np.add.at(diag, nd * conn + a, Kl[m.elem][:, np.arange(8) * nd + a, np.arange(8) * nd + a])
    np.testing.assert_allclose(m.jacobi_inv(), 1.0 / diag, rtol=1e-13)
    if min(m.nx, m.ny, m.nz) > 1:
        np.testing.assert_allclose(diag, A.diagonal(), rtol=1e-13)
    # null space: constants (thermal), translations (elastic)
    for a in range(nd):
        t = np.zeros(m.ndof)
        t[a::nd] = 1.0
        assert np.abs(m.apply(t)).max() < 1e-12


def test_fluid_operator_brute():
    rng = np.random.default_rng(10)
    img = np.where(rng.random((4, 5, 6)) < 0.7, 0, 255).astype(np.uint8)
    m = make(FLUID, img, [0.0])
    K, _ = local(m)
    A = B.assemble_fluid(m.elem, m.nx, m.ny, m.nz, K.ravel())
    v = rng.standard_normal(m.ndof)
    assert rel(m.apply(v), A @ v) < 1e-13
    assert abs(A - A.T).max() < 1e-13


@pytest.mark.parametrize("analysis", [THERMAL, ELASTIC])
def test_rhs_and_column_brute(analysis):
    img = three_phase((3, 4, 5), 11)
    props = [1.0, 10.0, 3.5] if analysis == THERMAL else [1.0, 0.3, 10.0, 0.2, 4.0, 0.25]
    m = make(analysis, img, props, (0, 100, 255))
    Kl, CBl = local(m)
    nd = 1 if analysis == THERMAL else 3
    x = np.random.default_rng(12).standard_normal(m.ndof)
    for d in range(m.cdim):
        b = B.rhs(m.elem, m.nx, m.ny, m.nz, Kl, nd, d)
        assert rel(m.rhs(d), b) < 1e-14
        col = B.effective_column(m.elem, m.nx, m.ny, m.nz, CBl, nd, d, x)
        np.testing.assert_allclose(m.update_c(d, x).reshape(m.cdim, m.cdim)[:, d], col, rtol=1e-12, atol=1e-14)


def test_fluid_rhs_and_column_brute():
    rng = np.random.default_rng(13)
    img = np.where(rng.random((4, 5, 6)) < 0.7, 0, 255).astype(np.uint8)
    m = O.Model(FLUID, (6, 5, 4), elem_of(img, (0, 255)), [0.0], voxel=0.5)
    conn = B.connectivity(m.nx, m.ny, m.nz)
    fk, ids, _, nvel, _ = B.fluid_maps(m.elem, m.nx, m.ny, m.nz)
    x = rng.standard_normal(m.ndof)
    for d in range(3):
        b = np.zeros(m.ndof)
        acc = np.zeros(3)
        for e in np.flatnonzero(m.elem == 0):
            for k in range(8):
                n = conn[e, k]
                if fk[n] == 255:
                    b[3 * ids[n] + d] += 0.125
                    acc += 0.125 * x[3 * ids[n]:3 * ids[n] + 3]
        np.testing.assert_allclose(m.rhs(d), b, atol=1e-15)
        np.testing.assert_allclose(m.update_c(d, x).reshape(3, 3)[:, d], acc / m.N * 0.25, rtol=1e-12)


# ---------------------------------------------------------------------------
# closed forms of the homogenised tensor
# ---------------------------------------------------------------------------
def homogenize(m, **kw):
    kw.setdefault("tol", 1e-12)
    C, info = m.homogenize(**kw)
    assert all(v[0] for v in info.values())
    return C


def test_homogeneous_thermal():
    m = make(THERMAL, np.zeros((5, 4, 6), np.uint8), [3.7], (0,))
    np.testing.assert_allclose(homogenize(m), 3.7 * np.eye(3), atol=1e-10)


def test_homogeneous_elastic():
    m = make(ELASTIC, np.zeros((4, 6, 5), np.uint8), [2.5, 0.27], (0,))
    np.testing.assert_allclose(homogenize(m), B.isotropic_C(2.5, 0.27), atol=1e-9)


@pytest.mark.parametrize("normal", [0, 1, 2])
def test_laminate_thermal(normal):
    """Layers normal to one axis: harmonic mean across, arithmetic mean along."""
    ks = np.array([1.0, 10.0, 3.5])
    seq = np.array([0, 1, 1, 2, 0, 2, 1])
    shape = [3, 4, 5]
    shape[2 - normal] = len(seq)
    idx = [None, None, None]
    idx[2 - normal] = slice(None)
    img = np.broadcast_to(np.array([0, 100, 255], np.uint8)[seq][tuple(idx)], shape).copy()
    C = homogenize(make(THERMAL, img, ks, (0, 100, 255)))
    f = np.bincount(seq, minlength=3) / len(seq)
    expect = np.eye(3) * (f @ ks)
    expect[normal, normal] = 1.0 / (f @ (1.0 / ks))
    np.testing.assert_allclose(C, expect, atol=1e-9)


@pytest.mark.parametrize("normal", [0, 1, 2])
def test_laminate_elastic(normal):
    mats = [(1.0, 0.3), (10.0, 0.2), (4.0, 0.45)]
    seq = np.array([0, 1, 2, 2, 1, 0])
    shape = [3, 3, 3]
    shape[2 - normal] = len(seq)
    idx = [None, None, None]
    idx[2 - normal] = slice(None)
    img = np.broadcast_to(np.array([0, 100, 255], np.uint8)[seq][tuple(idx)], shape).copy()
    C = homogenize(make(ELASTIC, img, [p for mt in mats for p in mt], (0, 100, 255)))
    f = np.bincount(seq, minlength=3) / len(seq)
    expect = B.laminate_C([B.isotropic_C(*mt) for mt in mats], f, normal)
    assert rel(C, expect) < 1e-9


VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def transform_C(C, A):
    T = np.zeros((3, 3, 3, 3))
    for I, (i, j) in enumerate(VOIGT):
        for J, (k, l) in enumerate(VOIGT):
            for a, b in {(i, j), (j, i)}:
                for c, d in {(k, l), (l, k)}:
                    T[a, b, c, d] = C[I, J]
    Tp = np.einsum("ia,jb,kc,ld,abcd->ijkl", A, A, A, A, T)
    return np.array([[Tp[i, j, k, l] for (k, l) in VOIGT] for (i, j) in VOIGT])


# image-space operations and the coordinate maps they induce
# (x = column, y = -row, z = -layer)
IMAGE_OPS = {
    "flip_x": (lambda a: a[:, :, ::-1], np.diag([-1.0, 1, 1])),
    "flip_y": (lambda a: a[:, ::-1, :], np.diag([1.0, -1, 1])),
    "flip_z": (lambda a: a[::-1, :, :], np.diag([1.0, 1, -1])),
    "swap_xy": (lambda a: a.transpose(0, 2, 1), np.array([[0.0, -1, 0], [-1, 0, 0], [0, 0, 1]])),
    "swap_xz": (lambda a: a.transpose(2, 1, 0), np.array([[0.0, 0, -1], [0, 1, 0], [-1, 0, 0]])),
    "roll": (lambda a: np.roll(a, (1, 2, 3), axis=(0, 1, 2)), np.eye(3)),
}


@pytest.mark.parametrize("op", list(IMAGE_OPS))
def test_image_symmetries_elastic(op):
    f, A = IMAGE_OPS[op]
    img = three_phase((4, 5, 3), 14)
    props = [1.0, 0.3, 10.0, 0.2, 4.0, 0.25]
    C = homogenize(make(ELASTIC, img, props, (0, 100, 255)))
    Cp = homogenize(make(ELASTIC, np.ascontiguousarray(f(img)), props, (0, 100, 255)))
    assert rel(C, C.T) < 1e-9
    assert rel(Cp, transform_C(C, A)) < 1e-9


@pytest.mark.parametrize("op", list(IMAGE_OPS))
def test_image_symmetries_thermal(op):
    f, A = IMAGE_OPS[op]
    img = three_phase((5, 3, 4), 15)
    C = homogenize(make(THERMAL, img, [1.0, 10.0, 3.5], (0, 100, 255)))
    Cp = homogenize(make(THERMAL, np.ascontiguousarray(f(img)), [1.0, 10.0, 3.5], (0, 100, 255)))
    assert rel(C, C.T) < 1e-9
    assert rel(Cp, A @ C @ A.T) < 1e-9


@pytest.mark.parametrize("op", ["flip_x", "swap_xy", "roll"])
def test_image_symmetries_fluid(op):
    f, A = IMAGE_OPS[op]
    rng = np.random.default_rng(16)
    img = np.where(rng.random((5, 6, 6)) < 0.75, 0, 255).astype(np.uint8)
    # Stokes frame has z along increasing layer: the z-reflection sign is moot
    # for the ops used here (they leave z alone)
    C = homogenize(make(FLUID, img, [0.0]), tol=1e-11, solver=1)
    Cp = homogenize(make(FLUID, np.ascontiguousarray(f(img)), [0.0]), tol=1e-11, solver=1)
    assert rel(Cp, A @ C @ A.T) < 1e-7


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("analysis", [THERMAL, ELASTIC])
def test_solvers_vs_dense(analysis):
    """CG and MINRES (Jacobi and unpreconditioned) against a dense
    least-squares solve of the assembled singular system."""
    img = three_phase((3, 4, 3), 17)
    props = [1.0, 10.0, 3.5] if analysis == THERMAL else [1.0, 0.3, 10.0, 0.2, 4.0, 0.25]
    m = make(analysis, img, props, (0, 100, 255))
    Kl, CBl = local(m)
    nd = 1 if analysis == THERMAL else 3
    A = B.assemble(m.elem, m.nx, m.ny, m.nz, Kl, nd).toarray()
    for d in range(m.cdim):
        b = m.rhs(d)
        xs = np.linalg.lstsq(A, b, rcond=None)[0]
        col = B.effective_column(m.elem, m.nx, m.ny, m.nz, CBl, nd, d, xs)
        for solver in (0, 1):
            for precond in (0, 1):
                x, ok, it, res = m.solve(d, solver=solver, precond=precond, tol=1e-13)
                assert ok and it <= 2 * m.ndof
                np.testing.assert_allclose(m.update_c(d, x).reshape(m.cdim, m.cdim)[:, d], col,
                                           atol=1e-9 * np.abs(col).max())


@pytest.mark.parametrize("solver", [0, 1], ids=["cg", "minres"])
def test_inf_rule_pinned_by_residual(solver):
    """The INF rule (pcg.cu:130-146, pminres.cu:188-200) restated in the
    oracle stops at the first iteration whose true residual b - A x,
    evaluated here through the operator, has max|r| / max|b| <= tol; the
    iteration before it is above tol.  A rule on the preconditioned residual
    or on the L2 measure would stop elsewhere."""
    img = three_phase((5, 6, 4), 23)
    m = make(ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], (0, 100, 255))
    tol = 1e-5
    for d in (0, 4):
        b = m.rhs(d)
        x, ok, it, res = m.solve_rule(d, 1, solver=solver, tol=tol, maxit=5000)
        assert ok and it > 2
        true = np.abs(b - m.apply(x)).max() / np.abs(b).max()
        assert true <= tol * (1 + 1e-6)
        assert abs(res - true) <= 1e-6 * true
        xp, okp, itp, resp = m.solve_rule(d, 1, solver=solver, tol=tol, maxit=it - 1)
        assert not okp and itp == it - 1
        assert np.abs(b - m.apply(xp)).max() / np.abs(b).max() > tol


@pytest.mark.parametrize("solver", [0, 1], ids=["cg", "minres"])
def test_error_rule_pinned_by_iterate_change(solver):
    """The error rule: |alpha| max|d| / max|x| is the relative change of the
    iterate in the last step, max|x_k - x_(k-1)| / max|x_k|, read off two
    runs that stop at k and k - 1 iterations; CG also stops on the L2 rule."""
    img = three_phase((5, 4, 6), 29)
    m = make(THERMAL, img, [1.0, 10.0, 3.5], (0, 100, 255))
    tol = 1e-3
    stopped_by_change = 0
    for d in range(3):
        x, ok, it, res = m.solve_rule(d, 2, solver=solver, tol=tol, maxit=5000)
        assert it > 2
        xp, _, itp, _ = m.solve_rule(d, 2, solver=solver, tol=tol, maxit=it - 1)
        xpp, _, _, _ = m.solve_rule(d, 2, solver=solver, tol=tol, maxit=it - 2)
        change = np.abs(x - xp).max() / np.abs(x).max()
        prev = np.abs(xp - xpp).max() / np.abs(xp).max()
        assert abs(res - change) <= 1e-9 * max(change, 1e-300)
        # met means the rule's own metric is within tol (pcg.cu:233-236)
        assert ok == (change <= tol)
        assert prev > tol
        if solver == 1:
            assert change <= tol
        else:
            # CG stops at the first iteration where the change is within tol
            # or the L2 rule holds (its own count k2)
            _, _, k2, _ = m.solve_rule(d, 0, solver=solver, tol=tol, maxit=5000)
            assert it <= k2 and (change <= tol or it == k2)
        stopped_by_change += change <= tol
    # (Jacobi-CG reaches the L2 rule first on these images; MINRES has no
    # L2 alternative)
    if solver == 1:
        assert stopped_by_change == 3


def test_solve_rule_l2_equals_solve():
    img = three_phase((4, 5, 3), 31)
    m = make(THERMAL, img, [1.0, 10.0, 3.5], (0, 100, 255))
    for solver in (0, 1):
        a = m.solve(1, solver=solver, tol=1e-8)
        b = m.solve_rule(1, 0, solver=solver, tol=1e-8)
        # (OpenMP dot products: equal to the order of their partial sums)
        np.testing.assert_allclose(a[0], b[0], rtol=0, atol=1e-12 * np.abs(a[0]).max())
        assert a[1] == b[1] and abs(a[2] - b[2]) <= 1


def test_fluid_solvers_vs_dense():
    rng = np.random.default_rng(18)
    img = np.where(rng.random((4, 4, 5)) < 0.75, 0, 255).astype(np.uint8)
    m = make(FLUID, img, [0.0])
    K, _ = local(m)
    A = B.assemble_fluid(m.elem, m.nx, m.ny, m.nz, K.ravel()).toarray()
    for d in range(3):
        b = m.rhs(d)
        xs = np.linalg.lstsq(A, b, rcond=None)[0]
        col = m.update_c(d, xs).reshape(3, 3)[:, d]
        for solver in (0, 1):
            x, ok, it, res = m.solve(d, solver=solver, tol=1e-12, maxit=20000)
            assert ok
            np.testing.assert_allclose(m.update_c(d, x).reshape(3, 3)[:, d], col, atol=1e-8 * np.abs(col).max())


def test_initial_guess_and_zero_rhs():
    img = three_phase((4, 4, 4), 19)
    m = make(THERMAL, img, [1.0, 10.0, 3.5], (0, 100, 255))
    x, ok, it, _ = m.solve(0, tol=1e-12)
    assert ok and it > 0
    x2, ok2, it2, _ = m.solve(0, tol=1e-6, x0=x)
    assert ok2 and it2 == 0
    np.testing.assert_array_equal(x2, x)
    # homogeneous image, shear load with no seam coupling: the right-hand side
    # of a single material is still non-zero, so check the x0 = exact path only
    mh = make(THERMAL, np.zeros((3, 3, 3), np.uint8), [2.0], (0,))
    xh, okh, ith, _ = mh.solve(1, tol=1e-12)
    assert okh


@pytest.mark.parametrize("dims,nnodedof", [((5, 4, 3), 1), ((2, 7, 1), 3), ((1, 1, 1), 4), ((6, 6, 6), 3)])
def test_node_dof_map(dims, nnodedof):
    """Node (r, c, l) of the (ny+1)(nx+1)(nz+1) grid -> DOF.  Checked two
    independent ways: the closed-form expression of assembleNodeDofMap_3D
    (femhmg_3D.c:4-17, written with node counts m_nx = nx+1, m_ny = ny+1),
    and the periodic identification node (r mod ny, c mod nx, l mod nz)
    with index r + c*ny + l*nx*ny (PERIODICNUM_3D, utils.h:82)."""
    nx, ny, nz = dims
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    out = np.zeros(NX * NY * NZ, dtype=np.uint32)
    O.lib().orc_node_dof_map(nx, ny, nz, nnodedof, O._p(out))
    n = np.arange(NX * NY * NZ, dtype=np.int64)
    i = n % (NX * NY)
    closed = ((i - i // NY - ny * ((i % NY) // ny)) % (nx * ny) + ((n // (NX * NY)) % nz) * nx * ny) * nnodedof
    np.testing.assert_array_equal(out, closed)
    r, c, l = n % NY, (n // NY) % NX, n // (NX * NY)
    np.testing.assert_array_equal(out, ((r % ny) + (c % nx) * ny + (l % nz) * nx * ny) * nnodedof)
    # interior nodes are numbered like elements; the last row/column/layer wraps to the first
    grid = out.reshape(NZ, NX, NY)
    np.testing.assert_array_equal(grid[-1], grid[0])
    np.testing.assert_array_equal(grid[:, -1], grid[:, 0])
    np.testing.assert_array_equal(grid[:, :, -1], grid[:, :, 0])
    assert np.unique(grid[:-1, :-1, :-1]).size == nx * ny * nz


@pytest.mark.parametrize("dims", [(8, 6, 4), (7, 5, 3), (2, 2, 2), (9, 4, 6)])
def test_coarse_elem_map(dims):
    """Coarse element map of the recursive initial guess: element (r, c, l)
    of the coarse mesh takes element (2r, 2c, 2l) of the fine one, with
    floor(n/2) coarse elements per axis (assemble_coarse_copy,
    femhmg.c:1844-1916).  Checked against NumPy slicing of the element map
    and against the element map of the image subsampled in file order."""
    nx, ny, nz = dims
    img = three_phase((nz, ny, nx), 11)
    em = elem_of(img, (0, 100, 255))
    out = np.zeros((nx // 2) * (ny // 2) * (nz // 2), dtype=np.uint16)
    O.lib().orc_coarse_elem_map(O._p(em), nx, ny, nz, O._p(out))
    grid = em.reshape(nz, nx, ny)
    np.testing.assert_array_equal(out, grid[:2 * (nz // 2):2, :2 * (nx // 2):2, :2 * (ny // 2):2].reshape(-1))
    sub = img[:2 * (nz // 2):2, :2 * (ny // 2):2, :2 * (nx // 2):2]
    np.testing.assert_array_equal(out, elem_of(sub, (0, 100, 255)))


def test_interpl2_linear_fields():
    """Coarse -> fine x0: nodal values doubled (the fine element is half the
    size, and the unknown is in element units), linear fields reproduced away
    from the periodic wrap (chfem/src/femhmg/cudapcg/kernels/image.cu:19-121)."""
    L = O.lib()
    rows, cols, lays = 6, 8, 4
    nyc, nxc, nzc = rows // 2, cols // 2, lays // 2
    i, j, k = np.meshgrid(np.arange(nyc), np.arange(nxc), np.arange(nzc), indexing="ij")
    coarse = np.zeros((nzc, nxc, nyc, 2))
    coarse[..., 0] = 5.0
    coarse[..., 1] = (0.3 * i + 0.7 * j - 0.4 * k).transpose(2, 1, 0)
    coarse = np.ascontiguousarray(coarse.reshape(-1))
    fine = np.zeros(rows * cols * lays * 2)
    L.orc_interpl2(O._p(coarse), rows, cols, lays, 2, O._p(fine))
    fine = fine.reshape(lays, cols, rows, 2)
    np.testing.assert_allclose(fine[..., 0], 10.0)
    ii, jj, kk = np.meshgrid(np.arange(rows), np.arange(cols), np.arange(lays), indexing="ij")
    lin = (0.3 * ii + 0.7 * jj - 0.4 * kk).transpose(2, 1, 0)
    inner = (np.arange(lays)[:, None, None] < lays - 1) & (np.arange(cols)[None, :, None] < cols - 1) & \
            (np.arange(rows)[None, None, :] < rows - 1)
    np.testing.assert_allclose(fine[..., 1][inner], lin[inner], atol=1e-14)


# ---------------------------------------------------------------------------
# scalar density field (chfem/src/femhmg/femhmg.c:1760-1835)
# ---------------------------------------------------------------------------
def sdf_scales(field_elem, props_first, elem, quantise=True):
    """Element scales as the reference builds them: value times the first
    property of the element's material, quantised to 16 bits between the
    extremes and mapped back (femhmg.c:1800-1824, femhmg_thermal_3D.c:265)."""
    E = field_elem * props_first[elem]
    if not quantise:
        return E
    emax, emin = E.max(), E.min()
    if emax == emin:
        emin = emax - 1.0
    d = np.floor(65535 * (E - emin) / (emax - emin)).astype(np.uint16)
    return d * (1.0 / 65535.0) * (emax - emin) + emin


@pytest.mark.parametrize("analysis", [THERMAL, ELASTIC])
def test_sdf_equals_material_properties(analysis):
    """A field holding each voxel's modulus, on unit-modulus materials, is
    the same model as the materials themselves."""
    img = three_phase((4, 5, 6), 31)
    if analysis == THERMAL:
        props = [1.0, 10.0, 3.5]
        unit = [1.0, 1.0, 1.0]
        first = np.array(props)
    else:
        props = [1.0, 0.3, 10.0, 0.2, 4.0, 0.25]
        unit = [1.0, 0.3, 1.0, 0.2, 1.0, 0.25]
        first = np.array(props[0::2])
    m = make(analysis, img, props, (0, 100, 255))
    ms = make(analysis, img, unit, (0, 100, 255))
    ms.set_element_scale(first[ms.elem])
    v = np.random.default_rng(32).standard_normal(m.ndof)
    assert rel(ms.apply(v), m.apply(v, strategy=1)) < 1e-13
    np.testing.assert_allclose(ms.jacobi_inv(), m.jacobi_inv(), rtol=1e-13)
    for d in range(m.cdim):
        assert rel(ms.rhs(d), m.rhs(d)) < 1e-13
        np.testing.assert_allclose(ms.update_c(d, v), m.update_c(d, v), rtol=1e-12, atol=1e-14)


def test_sdf_brute_and_scaling():
    img = three_phase((3, 4, 5), 33)
    m = make(ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], (0, 100, 255))
    esc = np.random.default_rng(34).uniform(0.5, 3.0, m.N)
    m.set_element_scale(esc)
    Kl, _ = local(m)
    # brute force with per-element scales
    import scipy.sparse as sp
    conn = B.connectivity(m.nx, m.ny, m.nz)
    ldof = (3 * conn[:, :, None] + np.arange(3)[None, None, :]).reshape(m.N, 24)
    vals = (Kl[m.elem] * esc[:, None, None]).reshape(m.N, -1).ravel()
    A = sp.csr_matrix((vals, (np.repeat(ldof, 24, axis=1).ravel(), np.tile(ldof, (1, 24)).ravel())),
                      shape=(m.ndof, m.ndof))
    v = np.random.default_rng(35).standard_normal(m.ndof)
    assert rel(m.apply(v), A @ v) < 1e-13
    # a uniform field scales the effective tensor
    m2 = make(ELASTIC, img, [1.0, 0.3, 10.0, 0.2, 4.0, 0.25], (0, 100, 255))
    C1 = homogenize(m2)
    m2.set_element_scale(np.full(m2.N, 2.5))
    C2 = homogenize(m2)
    assert rel(C2, 2.5 * C1) < 1e-9


@pytest.mark.slow
@pytest.mark.parametrize("case", [("thermal", 0, {87: 129.0, 182: 80.4}, [1.0], 3),
                                  ("elastic", 1, {87: 39.7, 182: 210.0}, None, 6)], ids=["thermal", "elastic"])
def test_oracle_golden_sdf(case):
    """The reference's SDF runs (test_chfem.c:63,66): unit-modulus materials
    plus a float32 field with each voxel's modulus (the reference's
    ggg40_100_{thermal,elastic}.bin holds exactly the per-phase values
    129.0/80.4 and 39.7/210.0 over ggg40_100.raw), against the same golden
    column as the plain run."""
    name, an, field_props, _, dim = case
    img = np.fromfile(os.path.join(GOLDEN, "ggg40_100.raw"), dtype=np.uint8).reshape(100, 100, 100)
    lut = np.zeros(256, np.float32)
    for k, v in field_props.items():
        lut[k] = v
    field = lut[img]                                   # file order, float32
    colors = (87, 182)
    elem = elem_of(img, colors)
    field_elem = np.ascontiguousarray(field.transpose(0, 2, 1)).reshape(-1).astype(np.float64)
    if an == THERMAL:
        esc = sdf_scales(field_elem, np.array([1.0, 1.0]), elem)
        m = O.Model(THERMAL, (100, 100, 100), np.zeros_like(elem), [1.0])
        d = 0
    else:
        esc = sdf_scales(field_elem, np.array([1.0, 1.0]), elem)
        m = O.Model(ELASTIC, (100, 100, 100), elem, [1.0, 0.2225, 1.0, 0.3])
        d = 1
    m.set_element_scale(esc)
    C, it = m.homogenize(directions=[d], tol=1e-6, maxit=1000, strategy=1)
    ref = np.fromfile(os.path.join(GOLDEN, f"{name}_ggg40_100_result.bin")).reshape(dim, dim)
    assert np.abs(C[:, d] - ref[:, d]).max() / np.abs(ref).max() < 1e-4
