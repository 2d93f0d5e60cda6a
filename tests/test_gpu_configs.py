"""The benchmark configurations at their full sizes on the GPU, against the
fixtures tests/golden/make_configs.py wrote from the CPU FP64 oracle:

* C1  32^3 Gaussian-random-field conductivity: FP64 and mixed precision,
      NBN and EBE, against the oracle's 3x3 tensor (tol 1e-10);
* C2  128^3 sphere elasticity: FP64 and mixed, NBN and EBE (matrix cores),
      against the oracle's 6x6 tensor (tol 1e-9);
* headline 256^3 sphere elasticity: FP64 against the oracle's X and YZ
      columns (tol 1e-8), and the benchmark's own method (mixed precision at
      tol 1e-5 from the zero guess and from a learned guess, autotuned tiles)
      against them within the metric's 1e-4;
* C3  256^3 Stokes permeability: the library's fluid keys, DOF numbering,
      DOF->node map and border keys hash to the oracle's; MINRES in FP64 and
      mixed precision give a symmetric tensor with a positive diagonal, the
      two within 1e-6;
* C4  64 uneven samples: batched against one-at-a-time, and 8 of them
      against the oracle;
* C5  512^3 sphere elasticity: a learned guess trained on a reduced set of
      64^3 solutions seeds mixed PCG to 1e-6; its tensor agrees with the zero
      guess's within 1e-5 in fewer iterations.

Every test first checks the SHA-256 of its seeded image against the fixture,
so a fixture applies only to the image it was made from.  Tolerances:
max |C - C*| / max |C*| (the reference harness's measure,
test/test_chfem.c:340-351) at most 1e-6 for solves at tol <= 1e-8.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from chfem_amd import synthetic
from tests import helpers as H
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

EL = {0: (1.0, 0.3), 255: (10.0, 0.2)}


def fixture(part):
    with open(os.path.join(GOLDEN, f"config_{part}.json")) as f:
        return json.load(f)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def relmax(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return np.abs(a - b).max() / np.abs(b).max()


def fixture_matrix(fx, dim):
    Cf = np.zeros((dim, dim))
    for k, v in fx["columns"].items():
        Cf[:, int(k)] = v["column"]
    return Cf


def solve_columns(res, dirs):
    cols, its = {}, {}
    for d in dirs:
        col, st, ok = res.solve(d)
        assert ok, (d, st.iterations, st.residual)
        cols[d], its[d] = col, st.iterations
    return cols, its


# ---------------------------------------------------------------------------
# C1
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("strategy", [0, 1], ids=["nbn", "ebe"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_c1_thermal_grf(precision, strategy):
    fx = fixture("c1")
    img = synthetic.grf((32, 32, 32), sigma=2.0, vf=0.4, seed=1)
    assert sha(img) == fx["image_sha256"]
    Cf = fixture_matrix(fx, 3)
    with H.Resident(img, H.THERMAL, {0: 1.0, 255: 10.0}, tol=1e-10, precision=precision, strategy=strategy) as r:
        assert sha(r.maps()[0]) == fx["keys_sha256"]
        cols, its = solve_columns(r, range(3))
    C = np.stack([cols[d] for d in range(3)], axis=1)
    assert relmax(C, Cf) < 1e-6
    if precision == 0:
        # same algorithm and stop rule as the oracle: counts agree to
        # reduction-order noise
        for d in range(3):
            assert abs(its[d] - fx["columns"][str(d)]["iterations"]) <= 2


# ---------------------------------------------------------------------------
# C2
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2_image():
    fx = fixture("c2")
    img = synthetic.spheres((128, 128, 128), rmin=6, rmax=12, fraction=0.3, seed=2)
    assert sha(img) == fx["image_sha256"]
    return img, fx


@pytest.mark.parametrize("strategy", [0, 1], ids=["nbn", "ebe"])
@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_c2_elastic_spheres(c2_image, precision, strategy):
    img, fx = c2_image
    Cf = fixture_matrix(fx, 6)
    with H.Resident(img, H.ELASTIC, EL, tol=1e-8, precision=precision, strategy=strategy) as r:
        if strategy == 0:
            assert sha(r.maps()[0]) == fx["keys_sha256"]
        cols, its = solve_columns(r, range(6))
    C = np.stack([cols[d] for d in range(6)], axis=1)
    assert relmax(C, Cf) < 1e-6
    assert relmax(C, C.T) < 1e-6
    print(f"C2 {['fp64', 'mixed'][precision]} {['nbn', 'ebe'][strategy]}: iterations "
          f"{[its[d] for d in range(6)]} (oracle at tol 1e-9: "
          f"{[fx['columns'][str(d)]['iterations'] for d in range(6)]})")


# ---------------------------------------------------------------------------
# headline 256^3
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def headline():
    fx = fixture("headline")
    img = synthetic.spheres((256, 256, 256), rmin=6, rmax=12, fraction=0.3, seed=3)
    assert sha(img) == fx["image_sha256"]
    return img, fx


def test_headline_fp64_vs_oracle(headline):
    img, fx = headline
    dirs = sorted(int(k) for k in fx["columns"])
    with H.Resident(img, H.ELASTIC, EL, tol=1e-8, precision=0) as r:
        assert sha(r.maps()[0]) == fx["keys_sha256"]
        cols, its = solve_columns(r, dirs)
    for d in dirs:
        ref = np.array(fx["columns"][str(d)]["column"])
        assert np.abs(cols[d] - ref).max() / np.abs(ref).max() < 1e-6, (d, cols[d], ref)


@pytest.mark.timeout(900)
def test_headline_benchmark_method_vs_oracle(headline):
    """The benchmark's method: all 6 load cases resident, mixed precision at
    tolerance 1e-5 with the autotuned tiles, from the zero guess and from a
    learned initial guess trained on a reduced set of 64^3 images of the same
    recipe; the columns the oracle solved within the metric's 1e-4."""
    from chfem_amd import guess
    from chfem_amd.guessbench import solve_resident
    img, fx = headline
    dirs = sorted(int(k) for k in fx["columns"])
    samples = [guess.solve_fluctuations(synthetic.spheres((64,) * 3, rmin=6, rmax=12, fraction=0.3, seed=1003 + k),
                                        EL, "elastic", tol=1e-6) for k in range(8)]
    net = guess.LearnedGuess("elastic", seed=3)
    net.train(samples, epochs=10, lr=2e-3, batch=4)
    for learned in (False, True):
        C, stats, _ = solve_resident(img, 1e-5, net if learned else None)
        for d in dirs:
            ref = np.array(fx["columns"][str(d)]["column"])
            err = np.abs(C[:, d] - ref).max() / np.abs(ref).max()
            print(f"headline mixed tol 1e-5 {'learned' if learned else 'zero'} guess, load {d}: "
                  f"{stats[d]['iterations']} iterations, error {err:.2e}")
            assert err < 1e-4
        if learned:
            assert all(s["x0_used"] == 1 for s in stats)


# ---------------------------------------------------------------------------
# C3
# ---------------------------------------------------------------------------
@pytest.mark.timeout(600)
def test_c3_stokes_maps_and_tensor():
    fx = fixture("c3")
    img, seed = synthetic.porous_spheres((256, 256, 256), rmin=6, rmax=12, porosity=0.35, seed=4)
    assert seed == fx["seed"] and sha(img) == fx["image_sha256"]
    Ks = {}
    for precision in (0, 1):
        with H.Resident(img, H.FLUID, {0: 0.0, 255: 0.0}, tol=1e-8, precision=precision) as r:
            if precision == 0:
                ne, nd, nv, nb = r.sizes()
                assert (nv, nb, nd) == (fx["nvel"], fx["nborder"], fx["ndof"])
                _, fk, ids, d2n, bk = r.maps()
                assert sha(fk) == fx["fluid_keys_sha256"]
                assert sha(ids) == fx["dof_id_sha256"]
                assert sha(d2n) == fx["dof2node_sha256"]
                assert sha(bk) == fx["border_keys_sha256"]
            cols, its = solve_columns(r, range(3))
        K = np.stack([cols[d] for d in range(3)], axis=1)
        print(f"C3 {'mixed' if precision else 'fp64'} MINRES iterations {[its[d] for d in range(3)]}, K = {K.tolist()}")
        assert np.all(np.diag(K) > 0)
        assert relmax(K, K.T) < 1e-6
        Ks[precision] = K
    assert relmax(Ks[1], Ks[0]) < 1e-6


# ---------------------------------------------------------------------------
# C4
# ---------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_c4_batched_uneven_samples():
    from chfem_amd import batch
    from chfem_amd.batchbench import make_samples, solve_one
    fx = fixture("c4")
    samples = make_samples(64, 48, 160)
    assert [sha(img) for _, img, _ in samples] == fx["image_sha256"]
    Cb = {}
    for prop in ("conductivity", "elasticity"):
        idx = [k for k, (p, _, _) in enumerate(samples) if p == prop]
        Cs, info = batch.batch_homogenize([(samples[k][1], samples[k][2]) for k in idx], prop, tol=1e-8, maxit=20000)
        assert info["converged"]
        Cb.update(zip(idx, Cs))
    for k, v in fx["solved"].items():
        err = relmax(Cb[int(k)], np.array(v["C"]))
        print(f"C4 sample {k} ({samples[int(k)][0]}, {samples[int(k)][1].shape}): batched vs oracle {err:.2e}")
        assert err < 1e-6
    # one at a time through the single-system solver, same precision and tolerance
    for k in range(0, 64, 7):
        prop, img, mp = samples[k]
        Cs, _, _ = solve_one(img, prop, mp, 1e-8, 20000)
        assert relmax(Cb[k], Cs) < 1e-6


# ---------------------------------------------------------------------------
# C5
# ---------------------------------------------------------------------------
@pytest.mark.timeout(1200)
def test_c5_learned_guess_512():
    from chfem_amd import guess
    from chfem_amd.guessbench import solve_resident, training_image
    fx = fixture("c5")
    samples = [guess.solve_fluctuations(training_image(k, 64), EL, "elastic", tol=1e-6) for k in range(32)]
    net = guess.LearnedGuess("elastic", seed=0)
    hist = net.train(samples, epochs=8, lr=2e-3, batch=4)
    del samples
    assert hist[-1] < hist[0]
    img = synthetic.spheres((512, 512, 512), rmin=6, rmax=12, fraction=0.3, seed=5)
    assert sha(img) == fx["image_sha256"]
    C0, s0, _ = solve_resident(img, 1e-6, None)
    C1, s1, _ = solve_resident(img, 1e-6, net)
    it0 = sum(s["iterations"] for s in s0)
    it1 = sum(s["iterations"] for s in s1)
    diff = relmax(C1, C0)
    print(f"C5 512^3: zero guess {it0} iterations, learned guess {it1} iterations "
          f"(x0 residuals {[round(s['x0_residual'], 4) for s in s1]}), tensors differ by {diff:.2e}")
    assert all(s["x0_used"] == 1 for s in s1)
    assert diff < 1e-5
    assert it1 < it0


# ---------------------------------------------------------------------------
# the measurement entry point of these configs, end to end at a small size
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("config,size", [("c2", 24), ("c3", 32)])
def test_configbench_small(config, size):
    """python -m chfem_amd.configbench runs every method of its config and
    prints one JSON line whose runs converged with the same effective
    tensor in FP64 and mixed precision (to the solver tolerance)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, "-m", "chfem_amd.configbench", "--config", config, "--size", str(size)],
                       cwd=root, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr[-2000:]
    out = json.loads(p.stdout.strip().splitlines()[-1])
    runs = out["runs"]
    assert len(runs) == (4 if config == "c2" else 2)
    for r in runs.values():
        assert all(r["converged"]) and r["seconds"] > 0
    cols = [np.array(r["C_columns"]) for r in runs.values()]
    for c in cols[1:]:
        assert relmax(c, cols[0]) < 1e-4
