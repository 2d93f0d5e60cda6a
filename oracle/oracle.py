"""ctypes loader for the CPU FP64 oracle (oracle/chfem_oracle.c).

TEST INFRASTRUCTURE: only tests/, ``__graft_entry__.smoke()`` and the CPU
baseline leg of ``bench.py`` import this module.  It is the checker, never the
thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

THERMAL, ELASTIC, FLUID = 0, 1, 2
TEXP = 6  # load case index of the thermal expansion analysis (elastic)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, sz, u32, dbl, i32 = C.c_void_p, C.c_size_t, C.c_uint32, C.c_double, C.c_int
        L.orc_create.restype = vp
        L.orc_create.argtypes = [i32, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp, vp, dbl]
        L.orc_destroy.argtypes = [vp]
        L.orc_ndof.restype = sz
        L.orc_ndof.argtypes = [vp]
        L.orc_nvel.restype = u32
        L.orc_nvel.argtypes = [vp]
        L.orc_nborder.restype = u32
        L.orc_nborder.argtypes = [vp]
        L.orc_npore.restype = sz
        L.orc_npore.argtypes = [vp]
        L.orc_get_keys.argtypes = [vp, vp]
        L.orc_get_fluid_maps.argtypes = [vp, vp, vp, vp, vp]
        L.orc_get_local.argtypes = [vp, vp, vp]
        L.orc_apply.argtypes = [vp, i32, vp, vp]
        L.orc_jacobi_inv.argtypes = [vp, vp]
        L.orc_rhs.argtypes = [vp, i32, vp]
        L.orc_updateC.argtypes = [vp, i32, vp, vp]
        L.orc_solve_rule.restype = i32
        L.orc_solve_rule.argtypes = [vp, i32, i32, i32, i32, dbl, C.c_uint, i32, sz, vp, vp,
                                     C.POINTER(C.c_uint), C.POINTER(C.c_double)]
        L.orc_solve.restype = i32
        L.orc_solve.argtypes = [vp, i32, i32, i32, i32, dbl, C.c_uint, vp, vp,
                                C.POINTER(C.c_uint), C.POINTER(dbl)]
        L.orc_image_to_elem_map.argtypes = [vp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp, vp]
        L.orc_node_dof_map.argtypes = [C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp]
        L.orc_dof_material_map.argtypes = [C.c_uint, C.c_uint, C.c_uint, vp, vp]
        L.orc_ke_thermal_unit.argtypes = [vp]
        L.orc_ke_elastic.argtypes = [dbl, dbl, i32, vp, vp]
        L.orc_ke_fluid.argtypes = [vp, vp]
        L.orc_thermal_pack40.argtypes = [vp, dbl, vp]
        L.orc_set_element_scale.argtypes = [vp, vp]
        L.orc_interpl2.argtypes = [vp, C.c_uint, C.c_uint, C.c_uint, C.c_uint, vp]
        L.orc_coarse_elem_map.argtypes = [vp, C.c_uint, C.c_uint, C.c_uint, vp]
        L.orc_set_alpha.argtypes = [vp, vp]
        L.orc_thermal_expansion.argtypes = [vp, vp, vp, vp]
        L.orc_fields.argtypes = [vp, i32, vp, vp]
        L.orc_fluid_fields.argtypes = [vp, vp, vp, vp]
        L.orc_num_threads.restype = i32
        L.orc_set_threads.argtypes = [i32]
        L.orc_wtime.restype = dbl
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def image_to_elem_map(img, props_keys, refinement=1):
    """img: uint8 array shaped (nz, ny, nx) in file order."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    nz, ny, nx = img.shape
    keys = np.zeros(256, dtype=np.uint8)
    for color, idx in props_keys.items():
        keys[color] = idx
    out = np.zeros(nx * ny * nz * refinement ** 3, dtype=np.uint16)
    lib().orc_image_to_elem_map(_p(img), nx, ny, nz, refinement, _p(keys), _p(out))
    return out, (nx * refinement, ny * refinement, nz * refinement)


class Model:
    """Oracle model: element map (material index per element, element order
    row + col*ny + layer*nx*ny) and material properties."""

    def __init__(self, analysis, dims, elem_map, props, voxel=1.0):
        self.analysis = analysis
        self.nx, self.ny, self.nz = dims
        self.elem = np.ascontiguousarray(elem_map, dtype=np.uint16)
        self.nmat = int(self.elem.max()) + 1 if analysis != FLUID else 2
        self.props = np.zeros(512)
        p = np.asarray(props, dtype=np.float64).ravel()
        self.props[:p.size] = p
        if analysis == THERMAL:
            self.nmat = p.size
        elif analysis == ELASTIC:
            self.nmat = p.size // 2
        self.h = lib().orc_create(analysis, self.nx, self.ny, self.nz, self.nmat,
                                  _p(self.elem), _p(self.props), voxel)
        self.ndof = lib().orc_ndof(self.h)
        self.N = self.nx * self.ny * self.nz
        self.cdim = 6 if analysis == ELASTIC else 3

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def nvel(self):
        return lib().orc_nvel(self.h)

    @property
    def nborder(self):
        return lib().orc_nborder(self.h)

    def set_element_scale(self, esc):
        """Scalar density field: per-element scale of the local matrices (None removes it)."""
        if esc is None:
            lib().orc_set_element_scale(self.h, None)
            return
        self._esc = np.ascontiguousarray(esc, dtype=np.float64)
        assert self._esc.size == self.N
        lib().orc_set_element_scale(self.h, _p(self._esc))

    def keys(self):
        out = np.zeros(self.N, dtype=np.uint16)
        lib().orc_get_keys(self.h, _p(out))
        return out

    def fluid_maps(self):
        fk = np.zeros(self.N, dtype=np.uint8)
        ids = np.zeros(self.N, dtype=np.uint32)
        d2n = np.zeros(self.nvel + self.nborder, dtype=np.uint32)
        bk = np.zeros(max(self.nborder, 1), dtype=np.uint8)
        lib().orc_get_fluid_maps(self.h, _p(fk), _p(ids), _p(d2n), _p(bk))
        return fk, ids, d2n, bk[:self.nborder]

    def local(self):
        kn = {THERMAL: 64, ELASTIC: 576, FLUID: 1024}[self.analysis]
        cn = {THERMAL: 24, ELASTIC: 144, FLUID: 80}[self.analysis]
        nm = 1 if self.analysis == FLUID else self.nmat
        K = np.zeros(kn * nm)
        CB = np.zeros(cn * nm)
        lib().orc_get_local(self.h, _p(K), _p(CB))
        return K, CB

    def apply(self, v, strategy=0):
        v = np.ascontiguousarray(v, dtype=np.float64)
        q = np.zeros(self.ndof)
        lib().orc_apply(self.h, strategy, _p(v), _p(q))
        return q

    def jacobi_inv(self):
        d = np.zeros(self.ndof)
        lib().orc_jacobi_inv(self.h, _p(d))
        return d

    def rhs(self, direction):
        b = np.zeros(self.ndof)
        lib().orc_rhs(self.h, direction, _p(b))
        return b

    def update_c(self, direction, x, C=None):
        if C is None:
            C = np.zeros(self.cdim * self.cdim)
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib().orc_updateC(self.h, direction, _p(x), _p(C))
        return C

    def solve(self, direction, solver=0, precond=1, strategy=0, tol=1e-6, maxit=10000, x0=None):
        x = np.zeros(self.ndof)
        it = C.c_uint(0)
        res = C.c_double(0)
        x0p = None
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            x0p = _p(x0)
        ok = lib().orc_solve(self.h, direction, solver, precond, strategy, tol, maxit, x0p, _p(x),
                             C.byref(it), C.byref(res))
        return x, bool(ok), it.value, res.value

    def solve_rule(self, direction, resnorm, solver=0, precond=1, strategy=0, tol=1e-6, maxit=10000, x0=None,
                   nhmg=0):
        """solve() with the stopping rule resnorm (0 L2, 1 INF, 2 error over
        the first nhmg entries, 0: all) as pcg.cu / pminres.cu apply it."""
        x = np.zeros(self.ndof)
        it = C.c_uint(0)
        res = C.c_double(0)
        x0p = None
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            x0p = _p(x0)
        ok = lib().orc_solve_rule(self.h, direction, solver, precond, strategy, tol, maxit, resnorm, nhmg, x0p,
                                  _p(x), C.byref(it), C.byref(res))
        return x, bool(ok), it.value, res.value

    def set_alpha(self, alpha):
        """Thermal expansion coefficient per material (elastic)."""
        self._alpha = np.zeros(256)
        a = np.asarray(alpha, dtype=np.float64).ravel()
        self._alpha[:a.size] = a
        lib().orc_set_alpha(self.h, _p(self._alpha))

    def thermal_expansion(self, x, C):
        """alpha* (Voigt xx, yy, zz, yz, xz, xy) from the solution of load case TEXP
        and the 6x6 effective matrix C."""
        out = np.zeros(6)
        x = np.ascontiguousarray(x, dtype=np.float64)
        Cm = np.ascontiguousarray(C, dtype=np.float64).reshape(36)
        lib().orc_thermal_expansion(self.h, _p(x), _p(Cm), _p(out))
        return out

    def fields(self, direction, x):
        """Nodal flux (thermal, 3 per node) or stress (elastic, 6 per node) as exported."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((self.N, 6 if self.analysis == ELASTIC else 3))
        lib().orc_fields(self.h, direction, _p(x), _p(out))
        return out

    def fluid_fields(self, x):
        """Stokes: (velocity N x 3, pressure N) as exported."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        vel = np.zeros((self.N, 3))
        pres = np.zeros(self.N)
        lib().orc_fluid_fields(self.h, _p(x), _p(vel), _p(pres))
        return vel, pres

    def homogenize(self, directions=None, **kw):
        """Per-direction RHS -> solve -> updateC, as hmgSolveHomogenization
        (chfem/src/femhmg/femhmg.c:848-1198)."""
        if directions is None:
            directions = range(self.cdim)
        C = np.zeros(self.cdim * self.cdim)
        info = {}
        for d in directions:
            x, ok, it, res = self.solve(d, **kw)
            self.update_c(d, x, C)
            info[d] = (ok, it, res)
        return C.reshape(self.cdim, self.cdim), info


def read_nf(path):
    """Minimal .nf reader (chfem/src/femhmg/femhmg.c:1456-1586)."""
    toks = open(path).read().split()
    d = {}
    i = 0
    while i < len(toks):
        t = toks[i]
        if t == "%type_of_analysis":
            d["analysis"] = int(toks[i + 1]); i += 2
        elif t == "%type_of_solver":
            d["solver"] = int(toks[i + 1]); i += 2
        elif t == "%voxel_size":
            d["voxel"] = float(toks[i + 1]); i += 2
        elif t == "%solver_tolerance":
            d["tol"] = float(toks[i + 1]); i += 2
        elif t == "%number_of_iterations":
            d["maxit"] = int(toks[i + 1]); i += 2
        elif t == "%image_dimensions":
            d["dims"] = tuple(int(x) for x in toks[i + 1:i + 4]); i += 4
        elif t == "%refinement":
            d["refinement"] = int(toks[i + 1]); i += 2
        elif t == "%number_of_materials":
            d["nmat"] = int(toks[i + 1]); i += 2
        elif t == "%properties_of_materials":
            i += 1
            mats = []
            per = {0: 2, 1: 3, 2: 1}[d["analysis"]]
            for _ in range(d["nmat"]):
                mats.append(tuple(float(x) for x in toks[i:i + per]))
                i += per
            d["materials"] = mats
        else:
            i += 1
    return d


def model_from_nf(nf_path, img):
    """Builds an oracle Model from a .nf description and a (nz, ny, nx) image."""
    d = read_nf(nf_path)
    keys = {int(m[0]): k for k, m in enumerate(d["materials"])}
    ref = d.get("refinement", 1)
    elem, dims = image_to_elem_map(img, keys, ref)
    a = d["analysis"]
    if a == THERMAL:
        props = [m[1] for m in d["materials"]]
    elif a == ELASTIC:
        props = [v for m in d["materials"] for v in m[1:3]]
    else:
        props = [0.0]
    voxel = d.get("voxel", 1.0) / ref
    return Model(a, dims, elem, props, voxel), d
