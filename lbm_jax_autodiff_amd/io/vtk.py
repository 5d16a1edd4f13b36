"""Field export for ParaView (jax_lbm/io/vtk.py:13-110).

The reference builds an unstructured hexahedral mesh with meshio.  Here the
lattice is written as what it is, a legacy-VTK STRUCTURED_POINTS grid with
one cell per lattice node and the same cell-data names (density,
velocity_x/y/z, velocity_magnitude, solid, porosity), binary, with no
dependency beyond numpy.  Density and velocity come from the device kernels
(compute_density / compute_velocity); only the finished fields are copied to
the host.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _field as F
from ..core.ops import compute_density, compute_velocity


def write_structured_points(filename: str, dims, fields: dict, title: str = "lbm_jax_autodiff_amd") -> None:
    """Write cell fields on an nx x ny x nz grid of unit cells (x fastest).

    ``fields`` maps a name to an array of shape (nx, ny, nz) (indexed
    [x, y, z]); values are written as big-endian float64, as legacy VTK
    binary requires.
    """
    if not str(filename).endswith(".vtk"):
        raise ValueError("export writes legacy VTK: the file name must end in .vtk")
    nx, ny, nz = (int(d) for d in dims)
    with open(filename, "wb") as fh:
        fh.write(b"# vtk DataFile Version 3.0\n")
        fh.write((title[:250] + "\n").encode())
        fh.write(b"BINARY\nDATASET STRUCTURED_POINTS\n")
        fh.write(f"DIMENSIONS {nx + 1} {ny + 1} {nz + 1}\nORIGIN 0 0 0\nSPACING 1 1 1\n".encode())
        fh.write(f"CELL_DATA {nx * ny * nz}\n".encode())
        for name, arr in fields.items():
            a = np.asarray(arr, dtype=np.float64)
            if a.shape != (nx, ny, nz):
                raise ValueError(f"field {name!r} has shape {a.shape}, expected {(nx, ny, nz)}")
            fh.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n".encode())
            # VTK cell order is x fastest: the [x, y, z] array transposed to [z][y][x]
            fh.write(np.ascontiguousarray(a.transpose(2, 1, 0)).astype(">f8").tobytes())
            fh.write(b"\n")


def export_vtk(filename: str, f: torch.Tensor, solid=None, porosity=None) -> None:
    """Density, velocity (components and magnitude) and the optional solid
    mask / porosity of a population field f[i, x, y, z] (jax_lbm/io/vtk.py:13-110)."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    with torch.no_grad():
        rho = compute_density(f)
        u = compute_velocity(f, rho)
        speed = torch.sqrt((u * u).sum(0))
        fields = {"density": rho, "velocity_x": u[0], "velocity_y": u[1], "velocity_z": u[2],
                  "velocity_magnitude": speed}
        if solid is not None:
            fields["solid"] = solid
        if porosity is not None:
            fields["porosity"] = porosity
        host = {k: v.detach().to("cpu", torch.float64).numpy() for k, v in fields.items()}
    write_structured_points(filename, (nx, ny, nz), host)
