"""End-to-end GPU homogenization through the reference's entry points
(hmg* C ABI and compute_property), against the oracle and the reference's own
golden results (chfem test/test_chfem.py:6-22, test/test_chfem.c:340-351)."""
import os

import numpy as np
import pytest

import chfem_amd
from chfem_amd import synthetic
from tests import helpers as H
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def relmax(C, R):
    return np.abs(np.asarray(C) - np.asarray(R)).max() / np.abs(np.asarray(R)).max()


def test_golden_thermal():
    img = chfem_amd.import_raw(os.path.join(GOLDEN, "ggg40_100.raw"), (100, 100, 100))
    C = chfem_amd.compute_property("conductivity", img, direction="x",
                                   nf_filepath=os.path.join(GOLDEN, "thermal_ggg40_100.nf"))
    ref = np.fromfile(os.path.join(GOLDEN, "thermal_ggg40_100_result.bin")).reshape(3, 3)
    np.testing.assert_almost_equal(C, ref, decimal=4)
    assert relmax(C, ref) < 1e-4


def test_golden_elastic():
    img = chfem_amd.import_raw(os.path.join(GOLDEN, "ggg40_100.raw"), (100, 100, 100))
    C = chfem_amd.compute_property("elasticity", img, direction="y",
                                   nf_filepath=os.path.join(GOLDEN, "elastic_ggg40_100.nf"))
    ref = np.fromfile(os.path.join(GOLDEN, "elastic_ggg40_100_result.bin")).reshape(6, 6)
    np.testing.assert_almost_equal(C, ref, decimal=3)
    assert relmax(C, ref) < 1e-4


@pytest.mark.parametrize("solver", ["cg", "minres"])
def test_golden_permeability(solver):
    img = chfem_amd.import_raw(os.path.join(GOLDEN, "fibers_100x100x10.raw"), (10, 100, 100))
    C = chfem_amd.compute_property("permeability", img, direction="z", solver=solver,
                                   nf_filepath=os.path.join(GOLDEN, "fluid_fibers_100x100x10.nf"))
    ref = np.fromfile(os.path.join(GOLDEN, "fluid_fibers_100x100x10_result.bin")).reshape(3, 3)
    np.testing.assert_almost_equal(C, ref, decimal=14)
    assert relmax(C, ref) < 1e-4


def test_elastic_all_directions_vs_oracle():
    img = synthetic.spheres((10, 12, 14), rmin=2, rmax=3, fraction=0.3, seed=2)
    props = {0: (1.0, 0.3), 255: (10.0, 0.2)}
    C = np.array(chfem_amd.compute_property("elasticity", img, mat_props=props, solver_tolerance=1e-9))
    m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2])
    Cr, _ = m.homogenize(tol=1e-9)
    assert relmax(C, Cr) < 1e-6
    assert relmax(C, C.T) < 1e-6


def test_thermal_mixed_precision_vs_oracle():
    img = synthetic.grf((16, 16, 16), sigma=2.0, vf=0.4, seed=1)
    C = np.array(chfem_amd.compute_property("conductivity", img, mat_props={0: 1.0, 255: 10.0},
                                            solver_tolerance=1e-8, precision="mixed"))
    m = H.model(H.THERMAL, img, [1.0, 10.0])
    Cr, _ = m.homogenize(tol=1e-10)
    assert relmax(C, Cr) < 1e-6


def test_rhs_and_updateC_device_vs_oracle():
    from chfem_amd.lib import lib, ptr
    img = synthetic.spheres((8, 9, 10), rmin=2, rmax=3, fraction=0.3, seed=3)
    chfem_amd.export_for_chfem("/tmp/chfem_rhs_test.raw", img, 1, {0: (1.0, 0.3), 255: (10.0, 0.2)},
                               export_raw=False)
    L = lib()
    assert L.hmgInit(b"/tmp/chfem_rhs_test.nf", None, None, ptr(np.ascontiguousarray(img).reshape(-1))) == 1
    m = H.model(H.ELASTIC, img, [1.0, 0.3, 10.0, 0.2])
    try:
        rng = np.random.default_rng(9)
        for d in range(6):
            b = np.zeros(m.ndof)
            assert L.hmgAssembleRHS(d, ptr(b)) == 1
            br = m.rhs(d)
            assert np.abs(b - br).max() <= 1e-12 * np.abs(br).max()
            x = rng.standard_normal(m.ndof)
            C = np.zeros(36)
            assert L.hmgUpdateC(d, ptr(x), ptr(C)) == 1
            Cr = m.update_c(d, x)
            assert np.abs(C.reshape(6, 6)[:, d] - Cr.reshape(6, 6)[:, d]).max() <= 1e-12 * np.abs(Cr).max()
    finally:
        L.hmgEnd()


def write_nf(path, analysis, props, dims, tol=1e-6, maxit=1000, data_type="float32"):
    """.nf text (chfem/io.py:208-335 keyword layout) with a %data_type entry."""
    lines = ["%type_of_analysis", str(analysis), "", "%type_of_solver", "0", "", "%type_of_rhs", "0", "",
             "%voxel_size", "1.0", "", "%solver_tolerance", str(tol), "", "%number_of_iterations", str(maxit), "",
             "%image_dimensions", " ".join(str(d) for d in dims), "", "%refinement", "1", "",
             "%number_of_materials", str(len(props)), "", "%properties_of_materials"]
    lines += [" ".join(str(v) for v in p) for p in props]
    lines += ["", "%data_type", data_type, ""]
    open(path, "w").write("\n".join(lines))


@pytest.mark.parametrize("case", [("thermal", 0, {87: 129.0, 182: 80.4}, [(87, 1.0), (182, 1.0)], 3, 0),
                                  ("elastic", 1, {87: 39.7, 182: 210.0}, [(87, 1.0, 0.2225), (182, 1.0, 0.3)], 6, 1)],
                         ids=["thermal", "elastic"])
def test_golden_sdf(case, tmp_path):
    """Scalar density field runs of the reference's tests (test_chfem.c:63,66):
    unit-modulus materials and a float32 field with each voxel's modulus
    (built here from ggg40_100.raw: the reference's .bin holds exactly these
    per-phase values), through hmgInit(nf, raw, sdf), against the golden
    column of the plain run."""
    from chfem_amd.lib import lib, ptr
    name, an, field_props, mats, dim, d = case
    raw = os.path.join(GOLDEN, "ggg40_100.raw")
    img = np.fromfile(raw, dtype=np.uint8)
    lut = np.zeros(256, np.float32)
    for k, v in field_props.items():
        lut[k] = v
    binf = str(tmp_path / "field.bin")
    lut[img].tofile(binf)
    nf = str(tmp_path / "model.nf")
    write_nf(nf, an, mats, (100, 100, 100))
    L = lib()
    assert L.hmgInit(nf.encode(), raw.encode(), binf.encode(), None) == 1
    try:
        L.hmgSetHomogenizationFlag(d)
        L.hmgSolveHomogenization()
        import ctypes
        Cp = ctypes.cast(L.hmgGetConstitutiveMtx(), ctypes.POINTER(ctypes.c_double))
        C = np.array([Cp[i] for i in range(dim * dim)]).reshape(dim, dim)
    finally:
        L.hmgEnd()
    ref = np.fromfile(os.path.join(GOLDEN, f"{name}_ggg40_100_result.bin")).reshape(dim, dim)
    assert np.abs(C[:, d] - ref[:, d]).max() / np.abs(ref).max() < 1e-4


@pytest.mark.parametrize("kernel", ["mfma", "atomic"])
@pytest.mark.parametrize("which", ["thermal", "elastic"])
def test_sdf_vs_oracle(which, kernel, monkeypatch):
    """Element-by-element operator, Jacobi, RHS and solve with a random 16-bit
    density field against the oracle with the same element scales."""
    from chfem_amd.lib import lib, ptr
    if kernel == "atomic":
        monkeypatch.setenv("CHFEM_EBE", "atomic")
    rng = np.random.default_rng(61)
    img = rng.choice(np.array([0, 255], dtype=np.uint8), size=(9, 17, 12))
    if which == "thermal":
        m = H.model(H.THERMAL, img, [1.0, 1.0])
    else:
        m = H.model(H.ELASTIC, img, [1.0, 0.3, 1.0, 0.2])
    dens = rng.integers(0, 65536, size=m.N).astype(np.uint16)
    fmin, fmax = 0.7, 12.0
    m.set_element_scale(dens * (1.0 / 65535.0) * (fmax - fmin) + fmin)
    s = H.GpuSolver(m, strategy=1, tol=1e-9)
    try:
        assert lib().cudapcgSetParametricDensityField(ptr(dens), fmin, fmax) == 1
        v = rng.standard_normal(m.ndof)
        assert relmax(s.apply(v), m.apply(v, strategy=1)) < 1e-12
        assert relmax(s.precond(v), m.jacobi_inv() * v) < 1e-13
        b = m.rhs(m.cdim - 1)
        x, ok, st = s.solve(b)
        xr, okr, itr, _ = m.solve(m.cdim - 1, tol=1e-9, strategy=1)
        assert ok and okr
        assert relmax(m.update_c(m.cdim - 1, x), m.update_c(m.cdim - 1, xr)) < 1e-7
    finally:
        s.close()
    # the FP32 operator of mixed precision (the matrix-core kernel's 4x4
    # block form scales each element's gathered values) on an
    # FP32-representable input
    s = H.GpuSolver(m, strategy=1, precision=1, tol=1e-9)
    try:
        assert lib().cudapcgSetParametricDensityField(ptr(dens), fmin, fmax) == 1
        v32 = rng.standard_normal(m.ndof).astype(np.float32)
        q_ref = m.apply(v32.astype(np.float64), strategy=1)
        assert np.abs(s.apply32(v32) - q_ref).max() < 1e-5 * np.abs(q_ref).max()
    finally:
        s.close()
