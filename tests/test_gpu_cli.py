"""The chfem_exec command line (chfem/src/main.c:33-298) run as a child
process on the reference's own inputs, as test/test_chfem.c:290-354 runs them:
the saved effective matrix (-b) against the golden results at 1e-4 of
max |C|, plus the option paths (EBE, MINRES, density field, x export/import,
field export, report, thermal expansion)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "chfem_amd", "chfem_exec")
RAW = os.path.join(GOLDEN, "ggg40_100.raw")


def run_cli(args, cwd):
    print("[chfem_exec]", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stdout[-3000:]
    return r.stdout


def rel_col(C, ref, d):
    return np.abs(C[:, d] - ref[:, d]).max() / np.abs(C).max()


def copy_inputs(tmp_path, nf_name, raw_name):
    nf = str(tmp_path / nf_name)
    shutil.copy(os.path.join(GOLDEN, nf_name), nf)
    return nf, os.path.join(GOLDEN, raw_name)


@pytest.mark.parametrize("case", [
    ("thermal_ggg40_100.nf", "ggg40_100.raw", "thermal_ggg40_100_result.bin", 3, 0, []),
    ("thermal_ggg40_100.nf", "ggg40_100.raw", "thermal_ggg40_100_result.bin", 3, 0, ["-p", "1"]),
    ("thermal_ggg40_100.nf", "ggg40_100.raw", "thermal_ggg40_100_result.bin", 3, 0, ["-prec", "1"]),
    ("elastic_ggg40_100.nf", "ggg40_100.raw", "elastic_ggg40_100_result.bin", 6, 1, []),
    ("fluid_fibers_100x100x10.nf", "fibers_100x100x10.raw", "fluid_fibers_100x100x10_result.bin", 3, 2,
     ["-s", "1"]),
], ids=["thermal", "thermal-ebe", "thermal-mixed", "elastic", "fluid-minres"])
def test_cli_golden(tmp_path, case):
    nf_name, raw_name, res, dim, d, extra = case
    nf, raw = copy_inputs(tmp_path, nf_name, raw_name)
    out = str(tmp_path / "C.bin")
    run_cli([nf, raw, "-d", str(d), "-b", out] + extra, str(tmp_path))
    C = np.fromfile(out).reshape(dim, dim)
    ref = np.fromfile(os.path.join(GOLDEN, res)).reshape(dim, dim)
    assert rel_col(C, ref, d) < 1e-4


def write_sdf_inputs(tmp_path, name):
    """Unit-modulus .nf and a float32 field holding each voxel's modulus of
    ggg40_100.raw (the reference's ggg40_100_<name>.bin carries exactly these
    per-phase values)."""
    from tests.test_gpu_homogenization import write_nf
    img = np.fromfile(RAW, dtype=np.uint8)
    lut = np.zeros(256, np.float32)
    if name == "thermal":
        lut[87], lut[182] = 129.0, 80.4
        mats, an = [(87, 1.0), (182, 1.0)], 0
    else:
        lut[87], lut[182] = 39.7, 210.0
        mats, an = [(87, 1.0, 0.2225), (182, 1.0, 0.3)], 1
    field = str(tmp_path / "field.bin")
    lut[img].tofile(field)
    nf = str(tmp_path / "sdf.nf")
    write_nf(nf, an, mats, (100, 100, 100))
    return nf, field


# The 3D cases of the reference's C harness (test/test_chfem.c:156-232):
# thermal X and elastic Y, each NBN, EBE and density-field EBE, with one
# level of recursive coarse initial guess (num_of_recursions_initguess = 1,
# i.e. -r 1), compared at 1e-4 of max |C| (test_chfem.c:340-351).
HARNESS = [
    ("thermal", 0, ["-p", "0"]), ("thermal", 0, ["-p", "1"]), ("thermal", 0, ["sdf"]),
    ("elastic", 1, ["-p", "0"]), ("elastic", 1, ["-p", "1"]), ("elastic", 1, ["sdf"]),
]


@pytest.mark.parametrize("case", HARNESS, ids=["thermal-nbn", "thermal-ebe", "thermal-sdf", "elastic-nbn",
                                               "elastic-ebe", "elastic-sdf"])
def test_cli_reference_harness_recursive_guess(tmp_path, case):
    """-r 1: each load case is first solved on the 2x coarser mesh
    (hmgFindInitialGuesses, femhmg.c:561-846), whose solution is projected
    and interpolated onto the fine mesh as its initial guess (image.cu:19-121).
    The density-field runs print the reference's warning and start from zero
    (femhmg.c:572-577)."""
    name, d, extra = case
    dim = 3 if name == "thermal" else 6
    out = str(tmp_path / "C.bin")
    if extra == ["sdf"]:
        nf, field = write_sdf_inputs(tmp_path, name)
        log = run_cli(["-i", nf, RAW, "-f", field, "-d", str(d), "-r", "1", "-b", out], str(tmp_path))
        assert "not available" in log and "coarse mesh" not in log
    else:
        nf, raw = copy_inputs(tmp_path, f"{name}_ggg40_100.nf", "ggg40_100.raw")
        log = run_cli([nf, raw, "-d", str(d), "-r", "1", "-b", out] + extra, str(tmp_path))
        assert "Computing initial guesses in coarse mesh 1" in log
        its = [int(l.split("Solver ran")[1].split()[0]) for l in log.splitlines() if "Solver ran" in l]
        # one coarse solve and one fine solve; the fine one starts from the coarse solution
        assert len(its) >= 1
    C = np.fromfile(out).reshape(dim, dim)
    ref = np.fromfile(os.path.join(GOLDEN, f"{name}_ggg40_100_result.bin")).reshape(dim, dim)
    assert rel_col(C, ref, d) < 1e-4


def test_cli_elastic_ebe_golden(tmp_path):
    """Elastic Y on the element-by-element operator (matrix cores), without
    the recursive guess, against the golden column."""
    nf, raw = copy_inputs(tmp_path, "elastic_ggg40_100.nf", "ggg40_100.raw")
    out = str(tmp_path / "C.bin")
    run_cli([nf, raw, "-d", "1", "-p", "1", "-b", out], str(tmp_path))
    ref = np.fromfile(os.path.join(GOLDEN, "elastic_ggg40_100_result.bin")).reshape(6, 6)
    assert rel_col(np.fromfile(out).reshape(6, 6), ref, 1) < 1e-4


def test_cli_density_field(tmp_path):
    """-f: the thermal density-field run of test_chfem.c:63 (unit materials,
    float32 field holding each voxel's conductivity)."""
    from tests.test_gpu_homogenization import write_nf
    img = np.fromfile(RAW, dtype=np.uint8)
    lut = np.zeros(256, np.float32)
    lut[87], lut[182] = 129.0, 80.4
    field = str(tmp_path / "field.bin")
    lut[img].tofile(field)
    nf = str(tmp_path / "sdf.nf")
    write_nf(nf, 0, [(87, 1.0), (182, 1.0)], (100, 100, 100))
    out = str(tmp_path / "C.bin")
    run_cli(["-i", nf, RAW, "-f", field, "-d", "0", "-b", out], str(tmp_path))
    ref = np.fromfile(os.path.join(GOLDEN, "thermal_ggg40_100_result.bin")).reshape(3, 3)
    assert rel_col(np.fromfile(out).reshape(3, 3), ref, 0) < 1e-4


def test_cli_export_import_x_fields_report(tmp_path):
    """-xo writes <nf>_X_<d>.bin; -xi <nf>_X starts each load case from it
    (femhmg.c:1053-1084), so the second run stops within a few iterations
    (its true residual against the first run's recurrence) with the same C;
    -e writes the solution and flux files; -m writes the report."""
    nf, raw = copy_inputs(tmp_path, "thermal_ggg40_100.nf", "ggg40_100.raw")
    base = nf[:-3]
    out1, out2 = str(tmp_path / "C1.bin"), str(tmp_path / "C2.bin")
    log1 = run_cli([nf, raw, "-xo", "-e", "-b", out1, "-m", str(tmp_path / "report.txt")], str(tmp_path))
    for d in range(3):
        x = np.fromfile(f"{base}_X_{d}.bin")
        assert x.size == 100 ** 3
        np.testing.assert_array_equal(np.fromfile(f"{base}_temperature_{d}.bin"), x)
        assert np.fromfile(f"{base}_flux_{d}.bin").size == 3 * 100 ** 3
    assert "Solver ran" in open(tmp_path / "report.txt").read()
    log2 = run_cli([nf, raw, "-xi", base + "_X", "-b", out2], str(tmp_path))
    C1, C2 = np.fromfile(out1), np.fromfile(out2)
    assert np.abs(C1 - C2).max() < 1e-6 * np.abs(C1).max()
    its1 = [int(l.split("Solver ran")[1].split()[0]) for l in log1.splitlines() if "Solver ran" in l]
    its2 = [int(l.split("Solver ran")[1].split()[0]) for l in log2.splitlines() if "Solver ran" in l]
    assert len(its1) == 3 and len(its2) == 3
    assert all(b <= 5 for b in its2) and all(a > 10 for a in its1), (its1, its2)


def test_cli_thermal_expansion(tmp_path):
    """An elastic .nf with %thermal_expansion: the binary result holds C then
    alpha* (hmgSaveConstitutiveMtx, femhmg.c:1353-1365); a homogeneous
    expansion coefficient comes back exactly."""
    import chfem_amd
    from chfem_amd import io as cio
    sub = np.fromfile(RAW, dtype=np.uint8).reshape(100, 100, 100)[:40, :40, :40]
    img = np.where(sub == 87, 0, 255).astype(np.uint8)  # compute_property wants a phase 0
    mp = {0: (39.7, 0.2225, 1e-5), 255: (210.0, 0.3, 1e-5)}
    nf = str(tmp_path / "texp.raw")
    cio.export_for_chfem(nf, img, 1, mp, voxel_size=1.0, solver_tolerance=1e-9, solver_maxiter=20000)
    out = str(tmp_path / "C.bin")
    run_cli([nf[:-4] + ".nf", nf, "-b", out], str(tmp_path))
    v = np.fromfile(out)
    assert v.size == 42
    np.testing.assert_allclose(v[36:], 1e-5 * np.array([1, 1, 1, 0, 0, 0]), atol=1e-13)
    rows = chfem_amd.compute_property("elasticity", img, mat_props=mp, solver_tolerance=1e-9, solver_maxiter=20000)
    assert np.abs(np.array(rows[:6]).ravel() - v[:36]).max() < 1e-12 * np.abs(v[:36]).max()
