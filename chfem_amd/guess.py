"""Learned initial guess for the homogenization solves.

A periodic 3D convolutional encoder-decoder (``chfem_amd/csrc/guess.cpp``,
BF16 matrix-core kernels in ``guess.hip``) maps the material properties of a
voxel image to the fluctuation field of every load case.  It is trained here
on solutions the library computes itself; installed with ``activate()`` it
seeds every thermal/elastic load case of ``compute_property`` /
``hmgSolveHomogenization`` with ``ramp + prediction``, and the solver falls
back to the zero guess when the prediction's initial residual is larger.

The reference has no counterpart: it seeds PCG with zeros, a coarse-grid
solution (``-r``) or a guess read from a file (chfem/src/femhmg/femhmg.c:1053-1084).
"""
import ctypes as C
import os
import tempfile

import numpy as np

from . import io as cio
from .lib import SolveStats, lib, ptr

PHYSICS = {"conductivity": 0, "thermal": 0, "elasticity": 1, "elastic": 1}
# load case -> (seam axis, jumping component): X, Y, Z, YZ, XZ, XY
LOADS = [(0, 0), (1, 1), (2, 2), (1, 2), (2, 0), (1, 0)]


def ramp(load, nd, dims):
    """Total field of a homogeneous medium for a load case, in the solver's
    DOF layout: the unit-gradient ramp with its jump on the periodic seam
    (X: right face of the last column; Y: top of row 0; Z: near face of
    layer 0), element units."""
    nx, ny, nz = dims
    i = np.arange(nx * ny * nz)
    r, c, l = i % ny, (i // ny) % nx, i // (nx * ny)
    axis, comp = LOADS[load]
    val = [c, (ny - r) % ny, (nz - l) % nz][axis].astype(np.float64)
    x = np.zeros(nd * i.size)
    x[(comp if nd == 3 else 0)::nd] = val
    return x


def element_map(img, mat_props):
    """Element material indices (order of ``mat_props``) in the solver's
    element order (row + col*ny + layer*nx*ny) from a (z, y, x) image."""
    lut = np.zeros(256, dtype=np.uint16)
    for k, color in enumerate(mat_props):
        lut[color] = k
    return np.ascontiguousarray(lut[img].transpose(0, 2, 1)).reshape(-1)


def props_vector(mat_props, physics):
    if physics == 1:
        return np.array([v for p in mat_props.values() for v in p[:2]], dtype=np.float64)
    return np.array([float(p if np.isscalar(p) else p[0]) for p in mat_props.values()], dtype=np.float64)


class LearnedGuess:
    """A network handle.  physics: 'elastic' (6 load cases x 3 components) or
    'thermal' (3 load cases)."""

    def __init__(self, physics="elastic", seed=0):
        self.physics = PHYSICS[physics]
        self.nd = 3 if self.physics == 1 else 1
        self.h = lib().chfemGuessCreate(self.physics, seed)
        if not self.h:
            raise RuntimeError("chfemGuessCreate failed")
        self.nout = lib().chfemGuessNumOutputs(self.h)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().chfemGuessDestroy(h)
            self.h = None

    @property
    def num_params(self):
        return int(lib().chfemGuessNumParams(self.h))

    def get_params(self):
        w = np.zeros(self.num_params, dtype=np.float32)
        assert lib().chfemGuessGetParams(self.h, ptr(w)) == 1
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {w.size}")
        assert lib().chfemGuessSetParams(self.h, ptr(w)) == 1

    def gradients(self):
        g = np.zeros(self.num_params, dtype=np.float32)
        assert lib().chfemGuessGetGradients(self.h, ptr(g)) == 1
        return g

    def zero_gradients(self):
        assert lib().chfemGuessZeroGradients(self.h) == 1

    def save(self, path):
        np.savez(path, params=self.get_params(), physics=self.physics)

    def load(self, path):
        with np.load(path, allow_pickle=False) as f:
            if int(f["physics"]) != self.physics:
                raise ValueError("parameter file is for another physics")
            self.set_params(f["params"])

    def predict_elem(self, elem, props, nmat, dims):
        nx, ny, nz = dims
        out = np.zeros((nx * ny * nz, self.nout), dtype=np.float32)
        elem = np.ascontiguousarray(elem, dtype=np.uint16)
        props = np.ascontiguousarray(props, dtype=np.float64)
        if lib().chfemGuessPredict(self.h, ptr(elem), ptr(props), nmat, nx, ny, nz, ptr(out)) != 1:
            raise RuntimeError("prediction failed (grid dimensions must be multiples of 8)")
        return out

    def predict(self, img, mat_props):
        """Fluctuation fields (N, nout) of a (z, y, x) image."""
        nz, ny, nx = img.shape
        return self.predict_elem(element_map(img, mat_props), props_vector(mat_props, self.physics),
                                 len(mat_props), (nx, ny, nz))

    def initial_guess(self, img, mat_props, load):
        """x0 of a load case in the solver's DOF layout: ramp + prediction."""
        nz, ny, nx = img.shape
        y = self.predict(img, mat_props)
        x = ramp(load, self.nd, (nx, ny, nz))
        for a in range(self.nd):
            x[a::self.nd] += y[:, load * self.nd + a]
        return x

    def accumulate(self, sample):
        """Forward + backward on one training sample; gradients are summed
        into the network.  Returns the sample's mean squared error."""
        elem, props, nmat, dims, target = sample
        loss = C.c_double(0)
        nx, ny, nz = dims
        if lib().chfemGuessAccumulate(self.h, ptr(elem), ptr(props), nmat, nx, ny, nz,
                                      ptr(np.ascontiguousarray(target, dtype=np.float32)), C.byref(loss)) != 1:
            raise RuntimeError("chfemGuessAccumulate failed")
        return loss.value

    def step(self, lr, grad_scale=1.0):
        assert lib().chfemGuessStep(self.h, lr, grad_scale) == 1

    def activate(self):
        """Seeds subsequent thermal/elastic homogenizations with this network."""
        assert lib().hmgSetInitialGuessNetwork(self.h) == 1

    @staticmethod
    def deactivate():
        lib().hmgSetInitialGuessNetwork(None)

    def train(self, samples, epochs=10, lr=1e-3, batch=4, seed=0, log=None):
        """Adam on the mean squared error of the fluctuation fields; returns
        the mean loss of each epoch."""
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            order = rng.permutation(len(samples))
            tot = 0.0
            for k in range(0, len(order), batch):
                idx = order[k:k + batch]
                for j in idx:
                    tot += self.accumulate(samples[j])
                self.step(lr, 1.0 / len(idx))
            hist.append(tot / len(samples))
            if log:
                log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.4e}")
        return hist


def solve_fluctuations(img, mat_props, physics="elastic", tol=1e-6, precision="mixed", maxit=20000):
    """Solves every load case of a (z, y, x) image with the library and returns
    a training sample (elem, props, nmat, dims, target): target[v, load*nd + a]
    is the zero-mean fluctuation of component a (solution minus ramp)."""
    ph = PHYSICS[physics]
    nd = 3 if ph == 1 else 1
    nloads = 6 if ph == 1 else 3
    nz, ny, nx = img.shape
    L = lib()
    d = tempfile.mkdtemp(prefix="chfem_guess_")
    nf = os.path.join(d, "sample.raw")
    cio.export_for_chfem(nf, img, ph, mat_props, voxel_size=1.0, export_raw=False, solver_tolerance=tol,
                         solver_maxiter=maxit)
    nf = nf[:-4] + ".nf"
    flat = np.ascontiguousarray(img).reshape(-1)
    if L.hmgInit(nf.encode(), None, None, ptr(flat)) != 1:
        raise RuntimeError("hmgInit failed")
    try:
        L.hmgSetPrecisionFlag(1 if precision == "mixed" else 0)
        L.hmgSetSolverTolerance(tol)
        if L.hmgPrepare() != 1:
            raise RuntimeError("hmgPrepare failed")
        N = nx * ny * nz
        target = np.zeros((N, nd * nloads), dtype=np.float32)
        x = np.zeros(nd * N)
        for load in range(nloads):
            L.hmgSolveLoadCase(load, None, ptr(x), None)
            u = x - ramp(load, nd, (nx, ny, nz))
            for a in range(nd):
                ua = u[a::nd]
                target[:, load * nd + a] = ua - ua.mean()
        L.hmgRelease()
    finally:
        L.hmgSetPrecisionFlag(0)
        L.hmgEnd()
        os.remove(nf)
        os.rmdir(d)
    return (element_map(img, mat_props), props_vector(mat_props, ph), len(mat_props), (nx, ny, nz), target)


def solve_stats(img, mat_props, physics="elastic", tol=1e-6, precision="fp64", net=None):
    """Per-load-case solver statistics of one homogenization, with the given
    network as initial guess (None: zero guess)."""
    ph = PHYSICS[physics]
    L = lib()
    d = tempfile.mkdtemp(prefix="chfem_guess_")
    nf = os.path.join(d, "img.raw")
    cio.export_for_chfem(nf, img, ph, mat_props, voxel_size=1.0, export_raw=False, solver_tolerance=tol,
                         solver_maxiter=20000)
    nf = nf[:-4] + ".nf"
    flat = np.ascontiguousarray(img).reshape(-1)
    if L.hmgInit(nf.encode(), None, None, ptr(flat)) != 1:
        raise RuntimeError("hmgInit failed")
    try:
        L.hmgSetPrecisionFlag(1 if precision == "mixed" else 0)
        L.hmgSetHomogenizationFlag(6)
        L.hmgSetInitialGuessNetwork(net.h if net is not None else None)
        Cm = np.zeros(36)
        if L.hmgPrepare() != 1:
            raise RuntimeError("hmgPrepare failed")
        L.hmgSolveResident(ptr(Cm))
        st = (SolveStats * 6)()
        cnt = C.c_uint(0)
        L.hmgGetSolveStats(st, C.byref(cnt))
        L.hmgRelease()
    finally:
        L.hmgSetInitialGuessNetwork(None)
        L.hmgSetPrecisionFlag(0)
        L.hmgEnd()
        os.remove(nf)
        os.rmdir(d)
    n = 6 if ph == 1 else 3
    return Cm[:n * n].reshape(n, n), [(st[i].iterations, st[i].x0_residual, st[i].x0_used) for i in range(cnt.value)]
