"""Oracle: thermal expansion load case and exported nodal fields, pinned by
closed forms (chfem/src/femhmg/physical_phenomena/femhmg_elastic_3D.c:464-508,
863-1073, 1800-2645; femhmg_thermal_3D.c:694-1105; femhmg_fluid_3D.c:312-360).

* Homogeneous medium: the thermal load vanishes and alpha* = alpha; the
  exported flux/stress is the material tensor's column of the load case at
  every node, seam nodes included.
* Two isotropic phases (any geometry): Levin's exact relation.  A uniform
  strain eps0 (1,1,1,0,0,0) with 3K1 (eps0 - a1) = 3K2 (eps0 - a2) = lam gives
  the same hydrostatic stress lam in both phases; it solves the discrete
  problem exactly (u = 0), so alpha* = eps0 (1,1,1,0,0,0) - lam C*^-1 (1,1,1,0,0,0).
* Any medium: the mean of the nodal field over the nodes is the C column
  (the corner average of the trilinear strain is its centre value), and for
  Stokes the exported velocity sums to N / voxel^2 times the permeability
  column.
"""
import numpy as np
import pytest

from chfem_amd import io as cio
from oracle import oracle as O

IOTA = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])


def two_phase(dims, seed, frac=0.4):
    rng = np.random.default_rng(seed)
    return (rng.random(int(np.prod(dims))) < frac).astype(np.uint16)


def test_thermal_expansion_homogeneous():
    m = O.Model(O.ELASTIC, (4, 5, 3), np.zeros(60, np.uint16), [2.0, 0.25])
    m.set_alpha([3e-5])
    assert np.abs(m.rhs(O.TEXP)).max() < 1e-18
    Cm, _ = m.homogenize(tol=1e-12)
    x, ok, it, res = m.solve(O.TEXP, tol=1e-12)
    a = m.thermal_expansion(x, Cm)
    np.testing.assert_allclose(a, 3e-5 * IOTA, atol=1e-18)


@pytest.mark.parametrize("seed,props,alpha", [
    (5, [1.0, 0.3, 7.0, 0.2], [2e-5, 7e-6]),
    (6, [3.0, 0.1, 0.5, 0.4], [-1e-6, 4e-5]),
])
def test_thermal_expansion_levin(seed, props, alpha):
    dims = (6, 5, 4)
    m = O.Model(O.ELASTIC, dims, two_phase(dims, seed), props)
    m.set_alpha(alpha)
    b = m.rhs(O.TEXP)
    assert abs(b.reshape(-1, 3).sum(0)).max() < 1e-15 * np.abs(b).max() * b.size  # consistent load
    Cm, _ = m.homogenize(tol=1e-13, maxit=20000)
    x, ok, it, res = m.solve(O.TEXP, tol=1e-13, maxit=20000)
    assert ok
    a = m.thermal_expansion(x, Cm)
    (E1, n1, E2, n2), (a1, a2) = props, alpha
    K1, K2 = E1 / (3 * (1 - 2 * n1)), E2 / (3 * (1 - 2 * n2))
    eps0 = (K1 * a1 - K2 * a2) / (K1 - K2)
    lam = 3 * K1 * (eps0 - a1)
    levin = eps0 * IOTA - lam * np.linalg.solve(Cm, IOTA)
    assert np.abs(a - levin).max() < 1e-10 * np.abs(levin).max()
    # the phases differ, so the composite is not the plain average
    assert np.abs(a[:3] - np.mean(alpha)).max() > 1e-3 * np.abs(levin).max()


@pytest.mark.parametrize("analysis", [O.THERMAL, O.ELASTIC])
def test_fields_homogeneous(analysis):
    dims = (4, 5, 3)
    props = [3.0] if analysis == O.THERMAL else [2.0, 0.25]
    m = O.Model(analysis, dims, np.zeros(60, np.uint16), props)
    Cm, _ = m.homogenize(tol=1e-12)
    for d in range(m.cdim):
        x, ok, it, res = m.solve(d, tol=1e-12)
        f = m.fields(d, x)
        assert np.abs(f - Cm[:, d][None, :]).max() < 1e-12 * np.abs(Cm).max()


@pytest.mark.parametrize("analysis", [O.THERMAL, O.ELASTIC])
def test_fields_mean_is_C_column(analysis):
    dims = (6, 5, 4)
    props = [1.0, 9.0] if analysis == O.THERMAL else [1.0, 0.3, 7.0, 0.2]
    m = O.Model(analysis, dims, two_phase(dims, 7), props)
    Cm, _ = m.homogenize(tol=1e-12, maxit=20000)
    for d in range(m.cdim):
        x, ok, it, res = m.solve(d, tol=1e-12, maxit=20000)
        f = m.fields(d, x)
        assert np.abs(f.mean(0) - Cm[:, d]).max() < 1e-12 * np.abs(Cm).max()
        # the field is not uniform: the seam and material terms both matter
        assert np.abs(f - f.mean(0)).max() > 1e-3 * np.abs(Cm).max()


def test_fluid_fields_sum_is_permeability():
    dims = (8, 7, 6)
    rng = np.random.default_rng(3)
    elem = (rng.random(int(np.prod(dims))) < 0.25).astype(np.uint16)
    m = O.Model(O.FLUID, dims, elem, [0.0], voxel=0.5)
    for d in range(3):
        x, ok, it, res = m.solve(d, solver=1, tol=1e-12, maxit=20000)
        Ccol = m.update_c(d, x).reshape(3, 3)[:, d]
        vel, pres = m.fluid_fields(x)
        fk, ids, _, _ = m.fluid_maps()
        assert np.all(vel[fk != 255] == 0) and np.all(pres[fk == 0] == 0)
        np.testing.assert_allclose(vel.sum(0) * 0.25 / m.N, Ccol, rtol=1e-12, atol=1e-18)
        np.testing.assert_array_equal(pres[fk > 0], x[3 * m.nvel + ids[fk > 0]])


def test_field_import_layout(tmp_path):
    """Exported fields are node by node (row fastest, then column, then layer)."""
    nz, ny, nx = 2, 3, 4
    n = nz * ny * nx
    for ncomp, fn in ((1, cio.import_scalar_field_from_chfem), (3, cio.import_vector_field_from_chfem),
                      (6, cio.import_stress_field_from_chfem)):
        v = np.arange(n * ncomp, dtype=np.float64)
        p = str(tmp_path / f"f{ncomp}.bin")
        v.tofile(p)
        a = fn(p, (nz, ny, nx))
        b = fn(p, (nz, ny, nx), rotate_domain=True)
        for (l, r, c) in ((0, 0, 0), (1, 2, 3), (1, 0, 2)):
            i = r + c * ny + l * nx * ny
            for k in range(ncomp):
                if ncomp == 1:
                    assert a[l, r, c] == v[i] and b[c, r, l] == v[i]
                else:
                    assert a[k, l, r, c] == v[ncomp * i + k] and b[k, c, r, l] == v[ncomp * i + k]
