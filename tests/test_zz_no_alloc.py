"""The allocation-free entry point hmgUpdateCNoAlloc against hmgUpdateC,
through the C ABI: two calls in a row on one GPU, the second on a smaller
model, each column equal to the allocating entry point's (same kernels and
reduction order; FP64 results within 1e-14 relative, far inside the 1e-6 the
effective tensor is held to).  A Stokes model exercises the velocity-sum
branch.  The test's device vectors come from the HIP runtime the library is
linked against (libamdhip64.so.7, through ctypes): a second HIP runtime in
the process, such as the one torch bundles, cannot open the GPU once the
library's runtime holds it."""
import ctypes as C

import numpy as np
import pytest

from chfem_amd import synthetic
from tests import helpers as H

pytestmark = pytest.mark.gpu


HIP_H2D, HIP_D2H = 1, 2


class DeviceArray:
    """FP64 device array from the library's own HIP runtime."""

    def __init__(self, host):
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.host.nbytes, HIP_H2D) == 0

    def get(self):
        out = np.empty_like(self.host)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, HIP_D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def column_pair(img, analysis, mp, dirs):
    """(no-alloc columns, hmgUpdateC columns) for the listed load cases of
    one model: solve each load case resident, take the column from the
    device copy of its solution with the no-alloc entry point, then after
    hmgRelease recompute it with hmgUpdateC from the host solution."""
    na, xs = {}, {}
    r = H.Resident(img, analysis, mp, tol=1e-8)
    try:
        for d in dirs:
            _, _, ok, x = r.solve(d, want_x=True)
            assert ok
            xs[d] = x
            xd = DeviceArray(x)
            cd = DeviceArray(np.full(r.dim, np.nan))
            print(f"hmgUpdateCNoAlloc: analysis {analysis}, grid {img.shape}, load {d}", flush=True)
            assert r.L.hmgUpdateCNoAlloc(d, xd.ptr, cd.ptr) == 1
            na[d] = cd.get()
            xd.free()
            cd.free()
        r.L.hmgRelease()
        r.prepared = False
        ref = {}
        for d in dirs:
            Cf = np.zeros(r.dim * r.dim)
            assert r.L.hmgUpdateC(d, xs[d].ctypes.data_as(C.c_void_p), Cf.ctypes.data_as(C.c_void_p)) == 1
            ref[d] = Cf.reshape(r.dim, r.dim)[:, d].copy()
    finally:
        r.close()
    return na, ref


def check(na, ref):
    for d in ref:
        scale = np.abs(ref[d]).max()
        err = np.abs(na[d] - ref[d]).max() / scale
        print(f"  load {d}: no-alloc {na[d]} vs hmgUpdateC {ref[d]}: {err:.1e} "
              f"({'bit-identical' if np.array_equal(na[d], ref[d]) else 'differs'})", flush=True)
        assert np.all(np.isfinite(na[d]))
        assert err <= 1e-14


def test_no_alloc_column_two_calls_second_smaller():
    big = synthetic.spheres((20, 24, 28), rmin=3, rmax=5, fraction=0.3, seed=11)
    small = synthetic.spheres((8, 10, 12), rmin=2, rmax=3, fraction=0.3, seed=12)
    el = {0: (1.0, 0.3), 255: (10.0, 0.2)}
    na, ref = column_pair(big, H.ELASTIC, el, [0, 4])
    check(na, ref)
    na, ref = column_pair(small, H.ELASTIC, el, [3])
    check(na, ref)
    # thermal, smaller again
    tiny = synthetic.grf((6, 8, 10), sigma=1.5, vf=0.4, seed=13)
    na, ref = column_pair(tiny, H.THERMAL, {0: 1.0, 255: 10.0}, [1])
    check(na, ref)


def test_no_alloc_column_stokes():
    img = synthetic.porous_spheres((16, 16, 16), rmin=2, rmax=4, porosity=0.45, seed=21)[0]
    na, ref = column_pair(img, H.FLUID, {0: 0.0, 255: 0.0}, [2])
    check(na, ref)


def test_no_alloc_rejects_without_resident_model():
    from chfem_amd.lib import lib
    L = lib()
    xd, cd = DeviceArray(np.zeros(8)), DeviceArray(np.zeros(6))
    assert L.hmgUpdateCNoAlloc(0, xd.ptr, cd.ptr) == 0
    xd.free()
    cd.free()
