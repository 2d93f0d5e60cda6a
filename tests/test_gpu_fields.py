"""GPU thermal expansion load case and field export through the reference's
entry points (hmg* C ABI, compute_property), against the oracle and the
closed forms of tests/test_oracle_fields.py."""
import ctypes as C
import os

import numpy as np
import pytest

import chfem_amd
from chfem_amd import io as cio
from chfem_amd import synthetic
from chfem_amd.lib import lib, ptr
from tests import helpers as H

pytestmark = pytest.mark.gpu

IOTA = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])


def relmax(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def run_hmg(tmp_path, img, prop, mat_props, tol, strategy=0, precision=0, fields=False, solver=None):
    """hmgInit -> flags -> hmgSolveHomogenization, as runAnalysis
    (chfem/src/main.c:72-152); returns (C, alpha or None, nf path without .nf)."""
    analysis = {"conductivity": 0, "elasticity": 1, "permeability": 2}[prop]
    nf = str(tmp_path / f"{prop}_{strategy}_{precision}.raw")
    st = {0: 0, 1: 0, 2: 1}[analysis] if solver is None else solver
    cio.export_for_chfem(nf, img, analysis, mat_props, voxel_size=1.0, export_raw=False, solver_type=st,
                         solver_tolerance=tol, solver_maxiter=20000)
    nf = nf[:-4] + ".nf"
    flat = np.ascontiguousarray(img).reshape(-1)
    L = lib()
    assert L.hmgInit(nf.encode(), None, None, ptr(flat)) == 1
    try:
        L.hmgSetParallelStrategyFlag(strategy)
        L.hmgSetPrecisionFlag(precision)
        L.hmgSetSolverFlag(st)
        L.hmgSaveFields(1 if fields else 0, 0)
        L.hmgSolveHomogenization()
        n = L.hmgGetConstitutiveMtxDim()
        Cm = np.ctypeslib.as_array(L.hmgGetConstitutiveMtx(), shape=(n,)).copy()
        na = L.hmgGetThermalExpansionDim()
        a = np.ctypeslib.as_array(L.hmgGetThermalExpansion(), shape=(na,)).copy() if na else None
    finally:
        L.hmgSetPrecisionFlag(0)
        L.hmgEnd()
    k = int(round(np.sqrt(n)))
    return Cm.reshape(k, k), a, nf[:-3]


def levin(Cm, props, alpha):
    (E1, n1), (E2, n2) = props
    a1, a2 = alpha
    K1, K2 = E1 / (3 * (1 - 2 * n1)), E2 / (3 * (1 - 2 * n2))
    eps0 = (K1 * a1 - K2 * a2) / (K1 - K2)
    lam = 3 * K1 * (eps0 - a1)
    return eps0 * IOTA - lam * np.linalg.solve(Cm, IOTA)


@pytest.mark.parametrize("strategy,precision", [(0, 0), (1, 0), (0, 1)])
def test_thermal_expansion_vs_oracle(tmp_path, strategy, precision):
    img = synthetic.spheres((9, 10, 11), rmin=2, rmax=3, fraction=0.35, seed=11)
    mp = {0: (1.0, 0.3, 2e-5), 255: (8.0, 0.2, 6e-6)}
    Cg, ag, _ = run_hmg(tmp_path, img, "elasticity", mp, tol=1e-9, strategy=strategy, precision=precision)
    assert ag is not None and ag.shape == (6,)
    m = H.model(H.ELASTIC, img, [1.0, 0.3, 8.0, 0.2])
    m.set_alpha([2e-5, 6e-6])
    Cr, _ = m.homogenize(tol=1e-11, maxit=20000)
    x, ok, it, res = m.solve(6, tol=1e-11, maxit=20000)
    ar = m.thermal_expansion(x, Cr)
    assert relmax(Cg, Cr) < 1e-6
    assert relmax(ag, ar) < 1e-6
    # Levin's exact two-phase relation on the GPU's own C
    assert relmax(ag, levin(Cg, [(1.0, 0.3), (8.0, 0.2)], [2e-5, 6e-6])) < 1e-6


def test_thermal_expansion_compute_property(tmp_path):
    """compute_property with (E, nu, alpha) returns the 6x6 matrix plus the
    alpha row (chfem/src/chfem_wrapper.c:43-87); a homogeneous image gives
    alpha back exactly."""
    img = np.zeros((5, 6, 7), np.uint8)
    img[1:3, 2:5, 1:6] = 255
    out = chfem_amd.compute_property("elasticity", img, mat_props={0: (2.0, 0.25, 1e-5), 255: (2.0, 0.25, 1e-5)},
                                     solver_tolerance=1e-10)
    assert len(out) == 7 and len(out[6]) == 6
    np.testing.assert_allclose(out[6], 1e-5 * IOTA, atol=1e-14)
    with pytest.raises(ValueError):
        chfem_amd.compute_property("elasticity", img, mat_props={0: (2.0, 0.25, 1e-5), 255: (2.0, 0.25)})
    with pytest.raises(ValueError):
        chfem_amd.compute_property("elasticity", img, mat_props={0: (2.0, 0.25, 1e-5), 255: (2.0, 0.25, 0.0)},
                                   direction="x")


@pytest.mark.parametrize("prop,strategy", [("conductivity", 0), ("elasticity", 0), ("elasticity", 1)])
def test_fields_vs_oracle(tmp_path, prop, strategy):
    """Exported solution and nodal flux/stress: the field of the exported
    solution matches the oracle's field of that same solution, and its mean is
    the C column."""
    img = synthetic.spheres((7, 8, 9), rmin=2, rmax=3, fraction=0.3, seed=12)
    if prop == "conductivity":
        mp, props, an, nf_ = {0: 1.0, 255: 6.0}, [1.0, 6.0], H.THERMAL, 3
    else:
        mp, props, an, nf_ = {0: (1.0, 0.3), 255: (6.0, 0.2)}, [1.0, 0.3, 6.0, 0.2], H.ELASTIC, 6
    Cg, _, base = run_hmg(tmp_path, img, prop, mp, tol=1e-9, strategy=strategy, fields=True)
    m = H.model(an, img, props)
    shape = img.shape
    for d in range(Cg.shape[0]):
        sol = "temperature" if an == H.THERMAL else "displacement"
        fld = "flux" if an == H.THERMAL else "stress"
        x = np.fromfile(f"{base}_{sol}_{d}.bin")
        assert x.size == m.ndof
        f = np.fromfile(f"{base}_{fld}_{d}.bin").reshape(-1, nf_)
        fr = m.fields(d, x)
        assert relmax(f, fr) < 1e-12
        assert relmax(f.mean(0), Cg[:, d]) < 1e-12
        # the import helpers read the same values in image order
        if an == H.THERMAL:
            v = cio.import_vector_field_from_chfem(f"{base}_{fld}_{d}.bin", shape)
        else:
            v = cio.import_stress_field_from_chfem(f"{base}_{fld}_{d}.bin", shape)
        assert v.shape == (nf_,) + shape
        np.testing.assert_array_equal(v[:, 2, 3, 4], f[3 + 4 * shape[1] + 2 * shape[1] * shape[2]])


def test_fields_homogeneous_uniform(tmp_path):
    img = np.zeros((4, 5, 6), np.uint8)
    img[0, 0, 0] = 255
    Cg, _, base = run_hmg(tmp_path, img, "elasticity", {0: (3.0, 0.3), 255: (3.0, 0.3)}, tol=1e-11, fields=True)
    for d in range(6):
        f = np.fromfile(f"{base}_stress_{d}.bin").reshape(-1, 6)
        assert np.abs(f - Cg[:, d][None, :]).max() < 1e-9 * np.abs(Cg).max()


def test_fluid_fields(tmp_path):
    img = synthetic.spheres((8, 10, 12), rmin=2, rmax=3, fraction=0.3, seed=4)  # solid spheres, pore 0
    Cg, _, base = run_hmg(tmp_path, img, "permeability", None, tol=1e-10, fields=True)
    m = H.model(H.FLUID, img, [0.0])
    fk, ids, _, _ = m.fluid_maps()
    for d in range(3):
        vel = np.fromfile(f"{base}_velocity_{d}.bin").reshape(-1, 3)
        pres = np.fromfile(f"{base}_pressure_{d}.bin")
        assert vel.shape[0] == m.N and pres.size == m.N
        assert np.all(vel[fk != 255] == 0) and np.all(pres[fk == 0] == 0)
        np.testing.assert_allclose(vel.sum(0) / m.N, Cg[:, d], rtol=1e-12, atol=1e-20)
        x, ok, it, res = m.solve(d, solver=1, tol=1e-11, maxit=20000)
        vr, pr = m.fluid_fields(x)
        assert relmax(vel, vr) < 1e-7
