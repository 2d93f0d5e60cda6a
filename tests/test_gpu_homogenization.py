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
