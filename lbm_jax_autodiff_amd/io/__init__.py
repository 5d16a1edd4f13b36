"""jax_lbm.io equivalents."""
from .vtk import export_vtk, write_structured_points  # noqa: F401
