"""Host-side checks that need no GPU: the C-ABI library exports every entry
point the public headers declare, the Python layer's argument validation
(chfem/compute_properties.py:58-118), the .nf writer/reader
(chfem/io.py:208-335) and the seeded synthetic microstructures."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import chfem_amd
from chfem_amd import io as cio
from chfem_amd import synthetic
from tests.conftest import ROOT

DECL = re.compile(r"^\s*(?:[A-Za-z_][\w]*\s*\**\s+)+\**\s*([A-Za-z_]\w*)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"typedef[^;]*;", "", text, flags=re.S)
        for m in DECL.finditer(text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from chfem_amd.lib import LIB_PATH
    assert os.path.exists(LIB_PATH), "libchfem.so not built (python -m chfem_amd.build)"
    lib = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) > 80
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_reference_entry_points_declared():
    """The drop-in set: every cudapcg*/hmg* entry of the reference headers
    (chfem/src/femhmg/cudapcg/cudapcg.h:19-62, chfem/src/femhmg/femhmg.h:6-36)."""
    names = declared_functions()
    ref = """cudapcgInit cudapcgEnd cudapcgSetModelConstructorFcn cudapcgBuildModel cudapcgNewModel
    cudapcgSetModel cudapcgSetNumTol cudapcgSetMaxIterations cudapcgSetXReduceMode cudapcgSetXReduceShift
    cudapcgSetXReduceScale cudapcgSetReduceStabFactor cudapcgSetSolver cudapcgSetPreconditioner
    cudapcgSetResNorm cudapcgAllocateArrays cudapcgFreeArrays cudapcgSetRHS cudapcgSetX0 cudapcgSetImage
    cudapcgSetParametricDensityField cudapcgSetPoreMap cudapcgSetPeriodic2DOFMap cudapcgSetDOF2PeriodicMap
    cudapcgSetLclMtxs cudapcgSolve cudapcgSetHeaderString cudapcgPrintSolverReport cudapcgPrintSolverReport2
    cudapcgPrintSolverMetrics cudapcgPrintSolverMetrics2 cudapcgGetNumIterations cudapcgGetMaxNumIterations
    cudapcgGetResidual hmgInit hmgEnd hmgSetXReduceFlag hmgSetParallelStrategyFlag hmgSetSolverFlag
    hmgSetPreConditionerFlag hmgSetStoppingCriteria hmgSetPoreMappingStrategy hmgSetHomogenizationFlag
    hmgFindInitialGuesses hmgExportX hmgImportX hmgSolveHomogenization hmgPrintModelData
    hmgPrintConstitutiveMtx hmgGetConstitutiveMtx hmgGetConstitutiveMtxDim hmgGetThermalExpansion
    hmgGetThermalExpansionDim hmgSaveConstitutiveMtx hmgSaveFields hmgKeepTrackOfAnalysisReport
    hmgPrintReport initDefaultInput runAnalysis""".split()
    assert not [n for n in ref if n not in names]


def test_ctypes_signatures_cover_declarations():
    from chfem_amd.lib import SIGNATURES
    bound = {s[0] for s in SIGNATURES}
    missing = declared_functions() - bound
    assert not missing, sorted(missing)


@pytest.mark.parametrize("kw,msg", [
    (dict(property="stiffness"), "Invalid property"),
    (dict(property="conductivity", direction="w", mat_props={0: 1.0, 255: 2.0}), "Invalid direction"),
    (dict(property="conductivity", solver="gmres", mat_props={0: 1.0, 255: 2.0}), "Invalid solver"),
    (dict(property="conductivity", solver="cg2", mat_props={0: 1.0, 255: 2.0}), "Two-vectors"),
    (dict(property="conductivity"), "Material properties must be provided"),
    (dict(property="elasticity", mat_props={0: (1, .3, 1e-5), 255: (2, .2)}, direction="all"),
     "Expected thermal expansion"),
    (dict(property="elasticity", mat_props={0: (1, .3, 1e-5), 255: (2, .2, 1e-5)}, direction="x"),
     "Must be \"all\""),
    (dict(property="conductivity", mat_props={0: 1.0}, precision="fp16"), "Invalid precision"),
    (dict(property="permeability", nf_filepath="/nonexistent/x.nf"), "does not exist"),
])
def test_compute_property_validation(kw, msg):
    img = np.zeros((4, 4, 4), np.uint8)
    img[1:3] = 255
    with pytest.raises(ValueError, match=msg):
        chfem_amd.compute_property(array=img, **kw)


def test_compute_property_rejects_bad_images():
    img = np.full((4, 4, 4), 255, np.uint8)
    with pytest.raises(ValueError, match="No fluid phase"):
        chfem_amd.compute_property("permeability", img)
    with pytest.raises(ValueError, match="uint8"):
        chfem_amd.compute_property("conductivity", img.astype(np.int32), mat_props={255: 1.0})


@pytest.mark.parametrize("analysis,props", [
    (0, {0: 1.0, 255: 10.0}),
    (1, {0: (1.0, 0.3), 255: (10.0, 0.2)}),
    (2, None),
])
def test_nf_roundtrip(tmp_path, analysis, props):
    img = np.zeros((3, 4, 5), np.uint8)
    img[:, 1:3] = 255
    fn = str(tmp_path / "m.raw")
    cio.export_for_chfem(fn, img, analysis, props, voxel_size=2.5e-6, solver_type=1, solver_tolerance=1e-7,
                         solver_maxiter=123, refinement=2)
    d = cio.import_nf(fn[:-4] + ".nf")
    assert d["type_of_analysis"] == analysis
    assert d["type_of_solver"] == 1
    assert d["voxel_size"] == pytest.approx(2.5e-6)
    assert d["solver_tolerance"] == pytest.approx(1e-7)
    assert d["number_of_iterations"] == 123
    assert tuple(d["image_dimensions"]) == img.shape
    assert d["refinement"] == 2
    raw = cio.import_raw(fn, img.shape)
    np.testing.assert_array_equal(raw, img)


def test_nf_golden_files_parse():
    from tests.conftest import GOLDEN
    d = cio.import_nf(os.path.join(GOLDEN, "fluid_fibers_100x100x10.nf"))
    assert d["type_of_analysis"] == 2 and d["refinement"] == 2
    d = cio.import_nf(os.path.join(GOLDEN, "elastic_ggg40_100.nf"))
    assert d["type_of_analysis"] == 1 and list(d["image_dimensions"]) == [100, 100, 100]


def test_synthetic_generators():
    g = synthetic.grf((32, 32, 32), sigma=2.0, vf=0.4, seed=1)
    assert g.dtype == np.uint8 and set(np.unique(g)) <= {0, 255}
    assert (g == 255).mean() == pytest.approx(0.4, abs=1 / g.size)
    np.testing.assert_array_equal(g, synthetic.grf((32, 32, 32), sigma=2.0, vf=0.4, seed=1))
    s, _ = synthetic.spheres((48, 48, 48), rmin=6, rmax=12, fraction=0.3, seed=2), None
    assert 0.3 <= (s == 255).mean() < 0.4
    p, seed = synthetic.porous_spheres((40, 40, 40), rmin=4, rmax=8, porosity=0.35, seed=4)
    assert (p == 0).mean() <= 0.35 + 0.05
    assert synthetic.percolates(p == 0)


@pytest.mark.parametrize("module", ["opbench", "batchbench", "guessbench", "membench", "configbench"])
def test_measurement_modules_parse_arguments(module):
    """The measurement entry points import and parse their arguments without
    a GPU (--help exits before any library call)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-m", f"chfem_amd.{module}", "--help"], cwd=root, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "usage" in p.stdout
