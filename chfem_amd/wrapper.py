"""``chfem_amd.wrapper.run`` -- same call as the reference's CPython extension
``chfem.wrapper.run`` (chfem/src/chfem_wrapper.c:7-89), bound with ctypes to
the C ABI entry ``chfem_run`` of libchfem.so."""
import ctypes as C

import numpy as np

from .lib import lib, ptr


def run(array, nf_file, analysis_type, direction, solver, precondition, output_fields):
    """Runs one homogenization analysis; returns the n x n effective matrix as a
    list of row lists (n = 3 thermal/permeability, 6 elasticity)."""
    if not isinstance(array, np.ndarray):
        raise TypeError("run() argument 1 must be numpy.ndarray")
    a = np.ascontiguousarray(array, dtype=np.uint8).reshape(-1)
    out = np.zeros(42, dtype=np.float64)
    n = C.c_uint(0)
    alpha = C.c_uint(0)
    rc = lib().chfem_run(ptr(a), a.size, nf_file.encode(), int(analysis_type), int(direction), int(solver),
                         int(bool(precondition)), int(output_fields), ptr(out), C.byref(n), C.byref(alpha))
    if rc != 0:
        raise SystemError("chfem analysis failed")
    nn = n.value
    rows = [[float(out[i * nn + j]) for j in range(nn)] for i in range(nn)]
    if alpha.value:
        rows.append([float(out[nn * nn + j]) for j in range(alpha.value)])
    return rows
