"""The library's node maps, bit for bit against the oracle: after hmgInit,
hmgGetMaps returns the packed NBN material keys (femhmg_3D.c:19-75) or, for
Stokes, the node fluid keys, the node -> DOF ids, the DOF -> node map and the
border keys (femhmg_fluid_3D.c:370-491).  These are the maps the solver
uploads; the DOF numbering is observable through exported solutions.

Inputs go through the three readers the library has: a NumPy array
(chfem.wrapper.run, femhmg.c:1711-1758) and .raw files with and without mesh
refinement (femhmg.c:1644-1709).  hmgInit builds the maps on the host, so
these tests run without a GPU."""
import ctypes as C
import os

import numpy as np
import pytest

from chfem_amd import io as cio
from chfem_amd import synthetic
from chfem_amd.lib import lib, ptr
from oracle import oracle as O
from tests.conftest import GOLDEN


def sizes(L):
    ne, nd, nv, nb = C.c_uint(), C.c_uint(), C.c_uint(), C.c_uint()
    assert L.hmgGetSizes(C.byref(ne), C.byref(nd), C.byref(nv), C.byref(nb)) == 1
    return ne.value, nd.value, nv.value, nb.value


def library_maps(nf, raw=None, img=None):
    """hmgInit on a .nf with a .raw file or a NumPy image; the maps it built."""
    L = lib()
    flat = None if img is None else np.ascontiguousarray(img, dtype=np.uint8).reshape(-1)
    ok = L.hmgInit(nf.encode(), None if raw is None else raw.encode(), None,
                   None if flat is None else ptr(flat))
    assert ok == 1
    try:
        ne, nd, nv, nb = sizes(L)
        keys = np.zeros(ne, np.uint16)
        fk = np.zeros(ne, np.uint8)
        ids = np.zeros(ne, np.uint32)
        d2n = np.zeros(max(nv + nb, 1), np.uint32)
        bk = np.zeros(max(nb, 1), np.uint8)
        assert L.hmgGetMaps(ptr(keys), ptr(fk), ptr(ids), ptr(d2n), ptr(bk)) == 1
    finally:
        L.hmgEnd()
    return {"nelem": ne, "ndof": nd, "nvel": nv, "nborder": nb, "keys": keys, "fk": fk, "ids": ids,
            "d2n": d2n[:nv + nb], "bk": bk[:nb]}


def check_against_oracle(got, m):
    assert got["nelem"] == m.N and got["ndof"] == m.ndof
    if m.analysis == O.FLUID:
        fk, ids, d2n, bk = m.fluid_maps()
        assert (got["nvel"], got["nborder"]) == (m.nvel, m.nborder)
        np.testing.assert_array_equal(got["fk"], fk)
        np.testing.assert_array_equal(got["ids"], ids)
        np.testing.assert_array_equal(got["d2n"], d2n)
        np.testing.assert_array_equal(got["bk"], bk)
    else:
        np.testing.assert_array_equal(got["keys"], m.keys())


def write_case(tmp_path, img, analysis, mp, name="case"):
    nf = str(tmp_path / f"{name}.raw")
    cio.export_for_chfem(nf, img, analysis, mp, voxel_size=1.0, export_raw=True)
    return nf[:-4] + ".nf", nf


@pytest.mark.parametrize("shape,nphase", [((5, 7, 9), 3), ((1, 3, 2), 2), ((8, 4, 6), 4), ((3, 3, 3), 1)])
@pytest.mark.parametrize("analysis", [0, 1])
@pytest.mark.parametrize("reader", ["numpy", "raw"])
def test_nbn_keys(tmp_path, shape, nphase, analysis, reader):
    """Packed keys of multi-material images (colours listed out of numeric
    order in the .nf, as export_for_chfem sorts them) on ragged grids."""
    rng = np.random.default_rng(sum(shape) + nphase)
    colors = np.array([0, 40, 130, 255][:nphase], dtype=np.uint8)
    img = rng.choice(colors, size=shape)
    if analysis == 0:
        mp = {int(c): float(1 + 3 * k) for k, c in enumerate(colors)}
        props = [mp[int(c)] for c in colors]
    else:
        mp = {int(c): (float(1 + 3 * k), 0.2 + 0.03 * k) for k, c in enumerate(colors)}
        props = [v for c in colors for v in mp[int(c)]]
    nf, raw = write_case(tmp_path, img, analysis, mp)
    got = library_maps(nf, raw=raw if reader == "raw" else None, img=img if reader == "numpy" else None)
    nz, ny, nx = shape
    m = O.Model(analysis, (nx, ny, nz), synthetic.elem_map(img, [int(c) for c in colors]), props)
    check_against_oracle(got, m)


@pytest.mark.parametrize("shape,seed", [((12, 20, 16), 4), ((5, 9, 7), 8), ((2, 3, 4), 9)])
@pytest.mark.parametrize("reader", ["numpy", "raw"])
def test_fluid_maps_random_porous(tmp_path, shape, seed, reader):
    """Fluid keys, DOF numbering (fully fluid nodes first in node order, then
    border nodes), DOF -> node map and border keys of random pore spaces."""
    rng = np.random.default_rng(seed)
    img = np.where(rng.random(shape) < 0.4, 255, 0).astype(np.uint8)
    img.flat[0] = 0
    nf, raw = write_case(tmp_path, img, 2, {0: 0.0, 255: 0.0})
    got = library_maps(nf, raw=raw if reader == "raw" else None, img=img if reader == "numpy" else None)
    nz, ny, nx = shape
    m = O.Model(O.FLUID, (nx, ny, nz), synthetic.elem_map(img, [0, 255]), [0.0])
    check_against_oracle(got, m)


@pytest.mark.parametrize("case", [
    ("thermal_ggg40_100.nf", "ggg40_100.raw", (100, 100, 100)),
    ("elastic_ggg40_100.nf", "ggg40_100.raw", (100, 100, 100)),
    ("fluid_fibers_100x100x10.nf", "fibers_100x100x10.raw", (10, 100, 100)),
], ids=["thermal-ggg40", "elastic-ggg40", "fluid-fibers-refinement2"])
def test_reference_inputs(case):
    """The reference's own test inputs through the .raw reader, including the
    fluid case with mesh refinement 2 (200 x 200 x 20 elements)."""
    nf_name, raw_name, shape = case
    nf, raw = os.path.join(GOLDEN, nf_name), os.path.join(GOLDEN, raw_name)
    got = library_maps(nf, raw=raw)
    img = np.fromfile(raw, dtype=np.uint8).reshape(shape)
    m, _ = O.model_from_nf(nf, img)
    check_against_oracle(got, m)
