"""CPU restatement of the learned-initial-guess network (chfem_amd/csrc/guess.cpp,
guess.hip) in NumPy float64: forward pass, mean-squared-error loss and the
gradients of every parameter.

TEST INFRASTRUCTURE: only tests/ import this module.  It checks the GPU
network; it is never the thing measured or shipped.

The reference has no counterpart (it seeds PCG with zeros, a coarse solution
or a file).  With ``quantize=True`` the tensors the GPU stores in BF16 are
rounded to BF16 at the same points (packed weights, input features, stored
activations, activation gradients fed to the matrix cores), so the two sides
differ only by summation order; ``quantize=False`` is plain float64 (used for
the finite-difference check).

Tensors are (nz, nx, ny, C): flattening gives the voxel order
v = row + col*ny + layer*nx*ny with channels last.
"""
import numpy as np

SAME, DOWN, UP = 0, 1, 2
# (mode, Ci, Co, input slot, output slot, relu, skip slot); slots as guess.cpp
A_IN, A_E0A, A_E0, A_E1A, A_E1, A_E2A, A_E2, A_E3A, A_E3, A_D2, A_D1, A_D0 = range(12)


def layers(cout):
    return [(SAME, 16, 16, A_IN, A_E0A, 1, None), (SAME, 16, 16, A_E0A, A_E0, 1, None),
            (DOWN, 16, 32, A_E0, A_E1A, 1, None), (SAME, 32, 32, A_E1A, A_E1, 1, None),
            (DOWN, 32, 48, A_E1, A_E2A, 1, None), (SAME, 48, 48, A_E2A, A_E2, 1, None),
            (DOWN, 48, 48, A_E2, A_E3A, 1, None), (SAME, 48, 48, A_E3A, A_E3, 1, None),
            (UP, 48, 48, A_E3, A_D2, 1, A_E2), (UP, 48, 32, A_D2, A_D1, 1, A_E1),
            (UP, 32, 16, A_D1, A_D0, 1, A_E0), (SAME, 16, cout, A_D0, None, 0, None)]


def taps(mode):
    return 8 if mode == DOWN else 27


def split_params(w, cout):
    """Parameter vector -> [(W[Co, taps, Ci], b[Co])] in layer order."""
    out, off = [], 0
    for mode, ci, co, *_ in layers(cout):
        n = co * taps(mode) * ci
        W = w[off:off + n].reshape(co, taps(mode), ci)
        b = w[off + n:off + n + co]
        out.append((W, b))
        off += n + co
    assert off == len(w)
    return out


def bf16(a):
    """Round to BF16 (nearest even) and back, as the GPU stores tensors."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _offsets():
    # tap t = (dl+1)*9 + (dc+1)*3 + (dr+1)
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shift(x, dl, dc, dr):
    """y[l, c, r] = x[l+dl, c+dc, r+dr] (periodic)."""
    return np.roll(x, shift=(-dl, -dc, -dr), axis=(0, 1, 2))


def _up(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def conv_fwd(mode, x, W):
    if mode == DOWN:
        out = 0
        for t in range(8):
            dr, dc, dl = t & 1, (t >> 1) & 1, t >> 2
            out = out + x[dl::2, dc::2, dr::2] @ W[:, t, :].T
        return out
    src = _up(x) if mode == UP else x
    out = 0
    for t, (dl, dc, dr) in enumerate(_offsets()):
        out = out + _shift(src, dl, dc, dr) @ W[:, t, :].T
    return out


def conv_bwd(mode, x, W, dz):
    """(dx, dW) of out = conv_fwd(mode, x, W) given dz = dL/dout."""
    dW = np.zeros_like(W)
    if mode == DOWN:
        dx = np.zeros_like(x)
        for t in range(8):
            dr, dc, dl = t & 1, (t >> 1) & 1, t >> 2
            xs = x[dl::2, dc::2, dr::2]
            dW[:, t, :] = np.einsum("lcro,lcri->oi", dz, xs)
            dx[dl::2, dc::2, dr::2] = dz @ W[:, t, :]
        return dx, dW
    src = _up(x) if mode == UP else x
    dsrc = np.zeros_like(src)
    for t, (dl, dc, dr) in enumerate(_offsets()):
        dW[:, t, :] = np.einsum("lcro,lcri->oi", dz, _shift(src, dl, dc, dr))
        dsrc += _shift(dz @ W[:, t, :], -dl, -dc, -dr)
    if mode == UP:
        nz, nx, ny, C = dsrc.shape
        dx = dsrc.reshape(nz // 2, 2, nx // 2, 2, ny // 2, 2, C).sum(axis=(1, 3, 5))
    else:
        dx = dsrc
    return dx, dW


def features(elem, props, nmat, physics, dims):
    """Input tensor (16 channels): log10(p / max p) and (nu - 0.25) * 4."""
    nx, ny, nz = dims
    props = np.asarray(props, dtype=np.float64)
    p = props[0:2 * nmat:2] if physics == 1 else props[:nmat]
    tab = np.zeros((nmat, 2))
    tab[:, 0] = np.log10(p / p.max())
    if physics == 1:
        tab[:, 1] = (props[1:2 * nmat:2] - 0.25) * 4.0
    tab = tab.astype(np.float32).astype(np.float64)
    x = np.zeros((nz, nx, ny, 16))
    x[..., :2] = tab[np.asarray(elem).reshape(nz, nx, ny)]
    return x


def forward(w, x, cout, quantize=True):
    q = bf16 if quantize else (lambda a: a)
    P = split_params(w, cout)
    act = {A_IN: q(x)}
    y = None
    for (mode, ci, co, si, so, relu, sk), (W, b) in zip(layers(cout), P):
        z = conv_fwd(mode, act[si], q(W)) + b
        if sk is not None:
            z = z + act[sk]
        if relu:
            z = np.maximum(z, 0)
        if so is None:
            y = z
        else:
            act[so] = q(z)
    return y, act


def loss_and_grad(w, x, target, nout, cout, quantize=True):
    """MSE over the nout real output channels and its gradient (same layout as w)."""
    q = bf16 if quantize else (lambda a: a)
    P = split_params(w, cout)
    y, act = forward(w, x, cout, quantize)
    V = y.shape[0] * y.shape[1] * y.shape[2]
    scale = 1.0 / (V * nout)
    d = y[..., :nout] - target
    loss = scale * np.sum(d * d)
    gy = np.zeros_like(y)
    gy[..., :nout] = 2.0 * scale * d
    grads = [None] * len(P)
    gact = {A_E0: None, A_E1: None, A_E2: None}
    L = layers(cout)
    for k in range(len(L) - 1, -1, -1):
        mode, ci, co, si, so, relu, sk = L[k]
        W, b = P[k]
        g = gy if so is None else gact[so]
        if relu:
            g = g * (act[so] > 0)
        if sk is not None:
            gact[sk] = g if gact[sk] is None else gact[sk] + g
        dz = q(g)
        dx, dW = conv_bwd(mode, act[si], q(W), dz)
        grads[k] = (dW, dz.sum(axis=(0, 1, 2)))
        if si != A_IN:
            gact[si] = dx if gact.get(si) is None else gact[si] + dx
    flat = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
    return loss, flat, y
