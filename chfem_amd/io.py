"""Input/output helpers with the reference's behaviour (chfem/io.py:3-335):
neutral-file (.nf) reading and writing, raw uint8 images and the exported
fields."""
import numpy as np


def import_nf(filename):
    """Reads a .nf file into a dict (chfem/io.py:3-71)."""
    if not filename.endswith(".nf"):
        raise ValueError("Filename does not end with '.nf'")
    with open(filename) as f:
        lines = [line.rstrip() for line in f]
    data = {}
    for key in ("type_of_analysis", "type_of_solver", "type_of_rhs", "number_of_iterations", "refinement",
                "number_of_materials"):
        data[key] = int(lines[lines.index(f"%{key}") + 1])
    data["voxel_size"] = float(lines[lines.index("%voxel_size") + 1])
    data["solver_tolerance"] = float(lines[lines.index("%solver_tolerance") + 1])
    nx, ny, nz = (int(v) for v in lines[lines.index("%image_dimensions") + 1].split())
    data["image_dimensions"] = (nz, ny, nx)
    ii = lines.index("%properties_of_materials")
    mats = []
    for jj in range(1, data["number_of_materials"] + 1):
        parts = lines[ii + jj].split()
        mats.append((int(parts[0]), *(float(p) for p in parts[1:])))
    data["properties_of_materials"] = mats
    data["volume_fraction"] = None
    if "%volume_fraction" in lines:
        data["volume_fraction"] = tuple(float(v) for v in lines[lines.index("%volume_fraction") + 1].split())
    data["data_type"] = None
    if "%data_type" in lines:
        data["data_type"] = lines[lines.index("%data_type") + 1]
    if "%thermal_expansion" in lines:
        ii = lines.index("%thermal_expansion")
        data["thermal_expansion"] = [float(lines[ii + jj]) for jj in range(1, data["number_of_materials"] + 1)]
    return data


def import_raw(filename, shape=None, dtype=np.uint8):
    """Reads a raw image; shape is (z, y, x) (chfem/io.py:73-109)."""
    if not filename.endswith(".raw"):
        raise ValueError("Filename does not end with '.raw'")
    with open(filename, "rb") as f:
        raw = f.read()
    if shape is not None:
        expected = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if len(raw) != expected:
            raise ValueError(f"File size ({len(raw)}) does not match expected size ({expected})")
    a = np.frombuffer(raw, dtype=dtype)
    return a.reshape(shape) if shape is not None else a


def _read_field(filename, domain_shape, ncomp, rotate_domain, dtype):
    """Exported fields are stored node by node in the solver's order
    (row fastest, then column, then layer), ncomp values per node."""
    data = np.fromfile(filename, dtype=dtype)
    if len(domain_shape) == 2:
        ny, nx = domain_shape
        a = data.reshape((nx, ny) + ((ncomp,) if ncomp > 1 else ()))   # [x, y, (comp)]
        if ncomp == 1:
            return a if rotate_domain else a.T
        return a.transpose(2, 0, 1) if rotate_domain else a.transpose(2, 1, 0)
    nz, ny, nx = domain_shape
    a = data.reshape((nz, nx, ny) + ((ncomp,) if ncomp > 1 else ()))  # [z, x, y, (comp)]
    if ncomp == 1:
        return a.transpose(1, 2, 0) if rotate_domain else a.transpose(0, 2, 1)
    return a.transpose(3, 1, 2, 0) if rotate_domain else a.transpose(3, 0, 2, 1)


def import_scalar_field_from_chfem(filename, domain_shape, rotate_domain=False, dtype=np.float64):
    """Scalar field (temperature, pressure) of an exported analysis, shaped like
    the image (z, y, x), or (x, y, z) with rotate_domain (chfem/io.py:111-140)."""
    return _read_field(filename, domain_shape, 1, rotate_domain, dtype)


def import_vector_field_from_chfem(filename, domain_shape, rotate_domain=False, dtype=np.float64):
    """Vector field (flux, displacement, velocity): (3, z, y, x), or (3, x, y, z)
    with rotate_domain; 2 components in 2D (chfem/io.py:142-173)."""
    return _read_field(filename, domain_shape, 2 if len(domain_shape) == 2 else 3, rotate_domain, dtype)


def import_stress_field_from_chfem(filename, domain_shape, rotate_domain=False, dtype=np.float64):
    """Stress field, Voigt order xx, yy, zz, yz, xz, xy: (6, z, y, x), or
    (6, x, y, z) with rotate_domain; 3 components in 2D (chfem/io.py:175-206)."""
    return _read_field(filename, domain_shape, 3 if len(domain_shape) == 2 else 6, rotate_domain, dtype)


def export_for_chfem(filename, array, analysis_type, mat_props=None, voxel_size=1, solver_type=0, rhs_type=0,
                     refinement=1, export_raw=True, export_nf=True, solver_tolerance=1.e-6, solver_maxiter=10000,
                     tmp_nf_file=None):
    """Writes the .nf (and optionally .raw) that describe an analysis of array
    (shape (z, y, x)); materials are listed in ascending colour order
    (chfem/io.py:208-335)."""
    domain = np.asarray(array).astype(np.uint8)
    ids, counts = np.unique(domain, return_counts=True)
    materials = dict(sorted(zip(ids.tolist(), counts.tolist())))
    if mat_props is not None:
        want, have = set(mat_props.keys()), set(materials.keys())
        if want != have:
            msg = []
            if have - want:
                msg.append(f"Missing material IDs in 'mat_props' that are present in the array: {have - want}.")
            if want - have:
                msg.append(f"Extra material IDs in 'mat_props' that are not present in the array: {want - have}.")
            raise ValueError(" ".join(msg))
    props, alpha = [], []
    if analysis_type == 0:
        props = [f"{k} {v}" for k, v in mat_props.items()]
    elif analysis_type == 1:
        for k, v in mat_props.items():
            props.append(f"{k} {v[0]} {v[1]}")
            if len(v) == 3:
                alpha.append(f"{v[2]}")
            elif alpha:
                alpha.append("0.0")
    elif analysis_type == 2:
        props = [f"{k}" for k in materials.keys()]
    else:
        raise ValueError("Invalid analysis type.")
    vf = np.array(list(materials.values())) * 100.0 / np.prod(domain.shape)
    if domain.ndim == 2:
        dims = [domain.shape[1], domain.shape[0], 0]
    else:
        dims = [domain.shape[2], domain.shape[1], domain.shape[0]]
    jdata = {"type_of_analysis": analysis_type, "type_of_solver": solver_type, "type_of_rhs": rhs_type,
             "voxel_size": voxel_size, "solver_tolerance": solver_tolerance, "number_of_iterations": solver_maxiter,
             "image_dimensions": dims, "refinement": refinement, "number_of_materials": len(materials),
             "properties_of_materials": props, "volume_fraction": list(np.around(vf, 2)), "data_type": "uint8"}
    if alpha:
        jdata["thermal_expansion"] = alpha
    if export_nf:
        out = []
        for k, v in jdata.items():
            out.append("%" + k)
            if k in ("properties_of_materials", "thermal_expansion"):
                out.extend(v)
            elif isinstance(v, list):
                out.append(" ".join(str(float(x)) if isinstance(x, (np.floating,)) else str(x) for x in v))
            else:
                out.append(str(v))
            out.append("")
        text = "\n".join(out).strip()
        if tmp_nf_file is not None:
            tmp_nf_file.write(text.encode("utf-8"))
            tmp_nf_file.flush()
            tmp_nf_file.seek(0)
        else:
            with open(filename[:len(filename) - 4] + ".nf", "w") as f:
                f.write(text)
    if export_raw:
        if filename[-4:] != ".raw":
            filename = filename + ".raw"
        domain.tofile(filename)
    return jdata
