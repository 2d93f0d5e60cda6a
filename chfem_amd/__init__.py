"""chfem on MI355X: matrix-free image-based FEM homogenization with
hand-written gfx950 HIP kernels behind chfem's C and Python interfaces."""
from . import io, wrapper  # noqa: F401
from .compute_properties import compute_property  # noqa: F401
from .io import export_for_chfem, import_nf, import_raw  # noqa: F401

__all__ = ["compute_property", "import_raw", "import_nf", "export_for_chfem", "io", "wrapper"]
