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
