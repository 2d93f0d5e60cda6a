"""Pins the CPU restatement of the learned-guess network (oracle/guess_ref.py)
before it is used to check the GPU kernels: its gradients against central
finite differences of its own loss (float64, no BF16 rounding), and its
convolutions against direct per-voxel sums."""
import numpy as np
import pytest

from oracle import guess_ref as G


def random_params(cout, seed, scale=0.2):
    rng = np.random.default_rng(seed)
    n = sum(co * G.taps(m) * ci + co for m, ci, co, *_ in G.layers(cout))
    return rng.standard_normal(n) * scale


@pytest.mark.parametrize("cout,nout", [(16, 3), (32, 18)])
def test_gradients_finite_differences(cout, nout):
    dims = (8, 8, 8)
    rng = np.random.default_rng(1)
    elem = rng.integers(0, 2, size=np.prod(dims)).astype(np.uint16)
    props = [1.0, 0.3, 10.0, 0.2] if nout == 18 else [1.0, 10.0]
    x = G.features(elem, props, 2, 1 if nout == 18 else 0, dims)
    # a few channels carry signal beyond the two features so that every
    # input channel of the first layer sees a gradient
    x[..., 2:5] = rng.standard_normal(x[..., 2:5].shape)
    w = random_params(cout, 2)
    target = rng.standard_normal((8, 8, 8, nout))
    loss, g, _ = G.loss_and_grad(w, x, target, nout, cout, quantize=False)
    # probe parameters in every layer (weights and biases)
    off = 0
    probes = []
    for mode, ci, co, *_ in G.layers(cout):
        n = co * G.taps(mode) * ci
        probes += [off + int(rng.integers(0, n)), off + n + int(rng.integers(0, co))]
        off += n + co
    h = 1e-6
    for p in probes:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        lp = G.loss_and_grad(wp, x, target, nout, cout, quantize=False)[0]
        lm = G.loss_and_grad(wm, x, target, nout, cout, quantize=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[p]) <= 1e-5 * max(abs(fd), 1e-3 * abs(loss)), (p, fd, g[p])


def test_convolutions_direct_sums():
    rng = np.random.default_rng(3)
    nz, nx, ny = 4, 6, 8
    x = rng.standard_normal((nz, nx, ny, 3))
    W = rng.standard_normal((2, 27, 3))
    y = G.conv_fwd(G.SAME, x, W)
    for (l, c, r) in [(0, 0, 0), (3, 5, 7), (1, 2, 3)]:
        s = np.zeros(2)
        for t in range(27):
            dl, dc, dr = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
            s += W[:, t, :] @ x[(l + dl) % nz, (c + dc) % nx, (r + dr) % ny]
        np.testing.assert_allclose(y[l, c, r], s, rtol=1e-12)
    Wd = rng.standard_normal((2, 8, 3))
    yd = G.conv_fwd(G.DOWN, x, Wd)
    assert yd.shape == (2, 3, 4, 2)
    l, c, r = 1, 2, 3
    s = sum(Wd[:, t, :] @ x[2 * l + (t >> 2), 2 * c + ((t >> 1) & 1), 2 * r + (t & 1)] for t in range(8))
    np.testing.assert_allclose(yd[l, c, r], s, rtol=1e-12)
    xc = rng.standard_normal((2, 3, 4, 3))
    yu = G.conv_fwd(G.UP, xc, W)
    assert yu.shape == (4, 6, 8, 2)
    l, c, r = 3, 0, 7
    s = np.zeros(2)
    for t in range(27):
        dl, dc, dr = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
        s += W[:, t, :] @ xc[((l + dl) % 4) // 2, ((c + dc) % 6) // 2, ((r + dr) % 8) // 2]
    np.testing.assert_allclose(yu[l, c, r], s, rtol=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 1e-30, 65504.0])
    b = G.bf16(v)
    # representable values survive; others land on the nearest BF16 value
    assert b[0] == 1.0 and b[1] == 1.0 and b[2] == 1.0078125
    assert abs(b[3] - v[3]) <= 2 ** -8 * abs(v[3])
    assert np.all(np.abs(b - v) <= 2 ** -8 * np.abs(v))
