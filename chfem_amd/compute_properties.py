"""``compute_property`` -- the reference's Python entry point
(chfem/compute_properties.py:7-140) on the MI355X library: same arguments,
validation and return value, plus one optional keyword, ``precision``."""
import os
import tempfile

import numpy as np

from .io import export_for_chfem
from .lib import lib
from .wrapper import run

PRECISIONS = {"fp64": 0, "mixed": 1}


def compute_property(property, array, mat_props=None, voxel_size=1e-6, solver=None, solver_tolerance=1e-6,
                     solver_maxiter=10000, precondition=True, type_of_rhs=0, refinement=1, direction="all",
                     output_fields=None, nf_filepath=None, precision="fp64"):
    """Effective conductivity, elasticity or permeability of a voxel image.

    :param property: 'conductivity', 'elasticity' or 'permeability'
    :param array: 3D uint8 numpy array (z, y, x) of phase ids
    :param mat_props: {phase: k} (conductivity), {phase: (E, nu)} or {phase: (E, nu, alpha)}
        (elasticity; alpha, the thermal expansion coefficient, adds the thermal
        expansion analysis and needs direction 'all'), None (permeability)
    :param solver: 'cg', 'minres', 'cg3', 'minres3', 'cg2', 'minres2' (default cg; minres for permeability)
    :param direction: 'x', 'y', 'z', 'yz', 'xz', 'xy' or 'all'
    :param precision: 'fp64' (default) or 'mixed' (FP32 inner Krylov solve with FP64
        iterative refinement to solver_tolerance)
    :param output_fields: path prefix: writes <prefix>.nf and, per load case d,
        <prefix>_{temperature,flux|displacement,stress|velocity,pressure}_<d>.bin
    :return: the effective matrix as a list of row lists, followed by the
        effective thermal expansion (xx, yy, zz, yz, xz, xy) when alpha is given
    """
    properties = {"conductivity": 0, "elasticity": 1, "permeability": 2}
    if property not in properties:
        raise ValueError(f"Invalid property: {property}")
    analysis_type = properties[property]
    directions = {"x": 0, "y": 1, "z": 2, "yz": 3, "xz": 4, "xy": 5, "all": 6}
    if direction not in directions:
        raise ValueError(f"Invalid direction: {direction}")
    direction_int = directions[direction]
    if precision not in PRECISIONS:
        raise ValueError(f"Invalid precision: {precision}")

    thermal_expansion = False
    if mat_props is not None and analysis_type == 1:
        thermal_expansion = any(len(p) == 3 for p in mat_props.values())
    if thermal_expansion:
        if direction_int < 6:
            raise ValueError(f"Invalid direction for thermal expansion analysis: {direction}. Must be \"all\".")
        for mat, p in mat_props.items():
            if len(p) < 3:
                raise ValueError(f"Invalid properties {p} for material {mat}. Expected thermal expansion coefficient.")

    output_fields_flag = 1 if output_fields is not None else 0
    if solver is None:
        solver = "minres" if property == "permeability" else "cg"
    solvers = {"cg": 0, "minres": 1, "cg3": 2, "minres3": 3, "cg2": 4, "minres2": 5}
    if solver not in solvers:
        raise ValueError(f"Invalid solver type: {solver}")
    if solver in ("cg2", "minres2") and (output_fields_flag or property != "permeability"):
        raise ValueError(f"Invalid solver type: {solver}. Two-vectors solvers do not produce fields to be exported, "
                         "nor are implemented for conductivy or elasticity analyses.")
    solver_type = solvers[solver]

    tmp = None
    if nf_filepath is not None:
        if not os.path.exists(nf_filepath):
            raise ValueError(f"The file {nf_filepath} does not exist.")
        nf_filename = nf_filepath
    else:
        if mat_props is None and property != "permeability":
            raise ValueError("Material properties must be provided, e.g. [(phase_id, cond),] for conductivity or "
                             "[(phase_id, young, poisson),] for elasticity.")
        if array.dtype != np.uint8:
            raise ValueError("Domain must be uint8 dtype.")
        if 0 not in np.unique(array):
            raise ValueError("No fluid phase found in the domain.")
        if output_fields is None:
            tmp = tempfile.NamedTemporaryFile(delete=False, suffix=".nf")
        else:
            path, _ = os.path.split(output_fields)
            if path == "" or os.path.exists(path):
                tmp = open(output_fields + ".nf", "wb")
            else:
                raise ValueError(f"Output fields directory {path} does not exist.")
        export_for_chfem(None, array, analysis_type=analysis_type, mat_props=mat_props, voxel_size=voxel_size,
                         solver_type=solver_type, rhs_type=type_of_rhs, refinement=refinement, export_raw=False,
                         export_nf=True, solver_tolerance=solver_tolerance, solver_maxiter=solver_maxiter,
                         tmp_nf_file=tmp)
        nf_filename = tmp.name

    flat = np.ascontiguousarray(np.reshape(array, (array.size,)))
    lib().hmgSetPrecisionFlag(PRECISIONS[precision])
    try:
        eff = run(flat, nf_filename, analysis_type, direction_int, solver_type, precondition, output_fields_flag)
    finally:
        lib().hmgSetPrecisionFlag(0)
        if tmp is not None:
            tmp.close()
            if nf_filepath is None and output_fields is None:
                os.remove(tmp.name)
    return eff
