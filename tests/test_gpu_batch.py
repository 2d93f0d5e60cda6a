"""Batched solver on the GPU: several samples of different sizes, every load
case, solved together; each effective matrix against the oracle and against
the single-system solver."""
import numpy as np
import pytest

from chfem_amd import batch, synthetic
from tests import helpers as H

pytestmark = pytest.mark.gpu


def relmax(C, R):
    return np.abs(np.asarray(C) - np.asarray(R)).max() / np.abs(np.asarray(R)).max()


SHAPES = [(12, 20, 16), (9, 7, 13), (17, 11, 6), (8, 8, 8)]


@pytest.mark.parametrize("prop", ["conductivity", "elasticity"])
def test_batch_vs_oracle(prop):
    samples = []
    for k, shape in enumerate(SHAPES):
        img = synthetic.spheres(shape, rmin=1, rmax=3, fraction=0.3, seed=70 + k)
        if prop == "conductivity":
            mp = {0: 1.0, 255: 20.0 if k % 2 else 5.0}
        else:
            mp = {0: (1.0, 0.3), 255: (10.0 if k % 2 else 4.0, 0.2)}
        samples.append((img, mp))
    Cs, info = batch.batch_homogenize(samples, prop, tol=1e-8)
    assert info["converged"], info
    for (img, mp), C in zip(samples, Cs):
        if prop == "conductivity":
            m = H.model(H.THERMAL, img, [mp[0], mp[255]])
        else:
            m = H.model(H.ELASTIC, img, [*mp[0], *mp[255]])
        Cr, _ = m.homogenize(tol=1e-10)
        assert relmax(C, Cr) < 1e-6
    # iterations of every system are stated and positive
    assert all(s[0] > 0 for st in info["stats"] for s in st)


def test_batch_homogeneous_and_tiny():
    """A homogeneous sample (its solution is the ramp, C = material tensor) and
    a 1-voxel-thick sample in the same batch."""
    from tests.brute import isotropic_C
    a = np.zeros((6, 5, 7), np.uint8)
    b = np.zeros((3, 9, 1), np.uint8)
    b[1, 2:5, 0] = 255
    b[0, 7, 0] = 255
    Cs, info = batch.batch_homogenize([(a, {0: (2.0, 0.25)}), (b, {0: (1.0, 0.3), 255: (3.0, 0.3)})],
                                      "elasticity", tol=1e-9)
    assert info["converged"]
    np.testing.assert_allclose(Cs[0], isotropic_C(2.0, 0.25), atol=1e-7)
    m = H.model(H.ELASTIC, b, [1.0, 0.3, 3.0, 0.3])
    Cr, _ = m.homogenize(tol=1e-11)
    assert relmax(Cs[1], Cr) < 1e-6


@pytest.mark.parametrize("form", ["2", "1"], ids=["single-kernel", "dir-fused"])
@pytest.mark.parametrize("ring", ["16", "4", "2"])
def test_batch_step_forms_and_rings(form, ring, monkeypatch):
    """The batch loop's two step forms (the single-kernel step with z
    recovered from the ring, the default; the direction-fused operator with a
    separate z update) and ring widths down to 2 (z stored instead of
    recovered) give the oracle's tensors."""
    monkeypatch.setenv("CHFEM_BATCH_FORM", form)
    monkeypatch.setenv("CHFEM_BATCH_RING", ring)
    samples = []
    for k, shape in enumerate(SHAPES[:3]):
        img = synthetic.spheres(shape, rmin=1, rmax=3, fraction=0.3, seed=80 + k)
        samples.append((img, {0: (1.0, 0.3), 255: (12.0, 0.2)}))
    Cs, info = batch.batch_homogenize(samples, "elasticity", tol=1e-8)
    assert info["converged"], info
    for (img, mp), C in zip(samples, Cs):
        Cr, _ = H.model(H.ELASTIC, img, [*mp[0], *mp[255]]).homogenize(tol=1e-10)
        assert relmax(C, Cr) < 1e-6


def test_batch_learned_guess():
    """Learned initial guesses in a batch: a network trained on a few
    solutions of the same recipe starts every load case below the zero
    guess's residual, the tensors still meet the oracle's, and the guessed
    batch takes fewer iterations."""
    from chfem_amd import guess
    el = {0: (1.0, 0.3), 255: (10.0, 0.2)}
    train = [guess.solve_fluctuations(synthetic.spheres((16, 16, 16), rmin=2, rmax=4, fraction=0.3, seed=200 + k),
                                      el, "elastic", tol=1e-7) for k in range(8)]
    net = guess.LearnedGuess("elastic", seed=1)
    net.train(train, epochs=25, lr=3e-3, batch=4)
    samples = [(synthetic.spheres(s, rmin=2, rmax=4, fraction=0.3, seed=300 + k), el)
               for k, s in enumerate([(16, 16, 16), (24, 16, 16)])]
    C0, i0 = batch.batch_homogenize(samples, "elasticity", tol=1e-8)
    C1, i1 = batch.batch_homogenize(samples, "elasticity", tol=1e-8, guess=net)
    assert i0["converged"] and i1["converged"]
    for (img, mp), Ca, Cb in zip(samples, C0, C1):
        Cr, _ = H.model(H.ELASTIC, img, [*mp[0], *mp[255]]).homogenize(tol=1e-10)
        assert relmax(Ca, Cr) < 1e-6 and relmax(Cb, Cr) < 1e-6
    used = [s[3] for st in i1["stats"] for s in st]
    res = [s[4] for st in i1["stats"] for s in st]
    it0 = sum(s[0] for st in i0["stats"] for s in st)
    it1 = sum(s[0] for st in i1["stats"] for s in st)
    print(f"batch zero guess {it0} iterations, learned guess {it1} (x0 residuals {np.round(res, 3).tolist()})")
    assert all(u == 1 for u in used) and max(res) < 1.0
    assert it1 < it0
