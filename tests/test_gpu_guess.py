"""Learned initial guess on the GPU: the BF16 matrix-core lane layout, the
network's forward pass and parameter gradients against the CPU restatement
(oracle/guess_ref.py, BF16 rounding at the same points), training, and what
the trained network gains over the zero initial guess."""
import numpy as np
import pytest

from chfem_amd import guess, synthetic
from chfem_amd.lib import lib, ptr
from oracle import guess_ref as G

pytestmark = pytest.mark.gpu


def to_bf16_bits(a):
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def test_mfma_bf16_layout():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((16, 16))
    B = rng.standard_normal((16, 16))
    a16, b16 = to_bf16_bits(A), to_bf16_bits(B)
    D = np.zeros((16, 16), dtype=np.float32)
    assert lib().chfemGuessMfmaCheck(ptr(a16), ptr(b16), ptr(D)) == 1
    ref = G.bf16(A) @ G.bf16(B)
    np.testing.assert_allclose(D, ref, rtol=1e-5, atol=1e-5)


def sample_input(dims, physics, seed):
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims
    elem = rng.integers(0, 2, size=nx * ny * nz).astype(np.uint16)
    props = np.array([1.0, 0.3, 10.0, 0.2]) if physics == 1 else np.array([1.0, 10.0])
    return elem, props


def random_params(net, seed, scale=0.15):
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal(net.num_params) * scale).astype(np.float32)
    cout = 32 if net.physics == 1 else 16
    # padded input channels of the first layer carry no signal
    W0, _ = G.split_params(w.astype(np.float64), cout)[0]
    W0[:, :, 2:] = 0
    w[:W0.size] = W0.ravel()
    return w


def rel_l2(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("physics,dims", [(1, (16, 8, 24)), (0, (8, 16, 8))])
def test_forward_matches_cpu(physics, dims):
    net = guess.LearnedGuess("elastic" if physics == 1 else "thermal", seed=1)
    w = random_params(net, 2)
    net.set_params(w)
    elem, props = sample_input(dims, physics, 3)
    y = net.predict_elem(elem, props, 2, dims)
    cout = 32 if physics == 1 else 16
    x = G.features(elem, props, 2, physics, dims)
    yr, _ = G.forward(w.astype(np.float64), x, cout, quantize=True)
    nx, ny, nz = dims
    yr = yr.reshape(nx * ny * nz, cout)[:, :net.nout]
    assert rel_l2(y, yr) < 1e-2
    assert np.abs(y - yr).max() < 3e-2 * np.abs(yr).max()


@pytest.mark.parametrize("physics", [1, 0])
def test_gradients_match_cpu(physics):
    dims = (8, 16, 16)
    net = guess.LearnedGuess("elastic" if physics == 1 else "thermal", seed=4)
    w = random_params(net, 5)
    net.set_params(w)
    elem, props = sample_input(dims, physics, 6)
    nx, ny, nz = dims
    rng = np.random.default_rng(7)
    target = rng.standard_normal((nx * ny * nz, net.nout)).astype(np.float32)
    net.zero_gradients()
    loss = net.accumulate((elem, props, 2, dims, target))
    g = net.gradients()
    cout = 32 if physics == 1 else 16
    x = G.features(elem, props, 2, physics, dims)
    lr, gr, _ = G.loss_and_grad(w.astype(np.float64), x, target.reshape(nz, nx, ny, net.nout).astype(np.float64),
                                net.nout, cout, quantize=True)
    assert abs(loss - lr) < 1e-3 * lr
    # every layer's weight and bias gradient
    off = 0
    for mode, ci, co, *_ in G.layers(cout):
        n = co * G.taps(mode) * ci
        for a, b in ((off, off + n), (off + n, off + n + co)):
            ref = gr[a:b]
            if np.linalg.norm(ref) == 0:
                continue
            assert rel_l2(g[a:b], ref) < 3e-2, (mode, ci, co, a, rel_l2(g[a:b], ref))
        off += n + co


def test_training_reduces_loss():
    dims = (16, 16, 16)
    img = synthetic.spheres((16, 16, 16), rmin=2, rmax=4, fraction=0.3, seed=11)
    mp = {0: (1.0, 0.3), 255: (10.0, 0.2)}
    sample = guess.solve_fluctuations(img, mp, "elastic", tol=1e-6)
    net = guess.LearnedGuess("elastic", seed=0)
    hist = net.train([sample], epochs=30, lr=2e-3, batch=1)
    assert hist[-1] < 0.5 * hist[0]


def test_learned_guess_gain():
    """Trained on seeded 32^3 sphere images, the network's guess (ramp +
    prediction) must start a held-out image's solves at a lower residual and
    finish them in fewer PCG iterations than the zero guess."""
    mp = {0: (1.0, 0.3), 255: (10.0, 0.2)}
    samples = [guess.solve_fluctuations(synthetic.spheres((32, 32, 32), rmin=3, rmax=6, fraction=0.3, seed=100 + s),
                                        mp, "elastic", tol=1e-6) for s in range(16)]
    net = guess.LearnedGuess("elastic", seed=0)
    hist = net.train(samples, epochs=12, lr=2e-3, batch=4)
    assert hist[-1] < hist[0]
    test_img = synthetic.spheres((32, 32, 32), rmin=3, rmax=6, fraction=0.3, seed=999)
    C0, st0 = guess.solve_stats(test_img, mp, "elastic", tol=1e-6, net=None)
    C1, st1 = guess.solve_stats(test_img, mp, "elastic", tol=1e-6, net=net)
    it0 = sum(s[0] for s in st0)
    it1 = sum(s[0] for s in st1)
    assert all(s[2] == 1 for s in st1), st1
    assert max(s[1] for s in st1) < 1.0
    assert it1 < it0, (it0, it1)
    assert np.abs(C1 - C0).max() < 1e-4 * np.abs(C0).max()
