"""Legacy-VTK field export (jax_lbm/io/vtk.py): the file parses back to the
fields written, in VTK's x-fastest cell order."""

import numpy as np
import pytest
import torch

from lbm_jax_autodiff_amd.io import export_vtk, write_structured_points


def _read(path):
    data = open(path, "rb").read()
    fields = {}
    head, _, rest = data.partition(b"CELL_DATA ")
    assert b"DATASET STRUCTURED_POINTS" in head
    dims = [int(v) for v in head.split(b"DIMENSIONS ")[1].split(b"\n")[0].split()]
    ncell, _, rest = rest.partition(b"\n")
    n = int(ncell)
    while rest.startswith(b"SCALARS "):
        line, _, rest = rest.partition(b"\n")
        name = line.split()[1].decode()
        _, _, rest = rest.partition(b"\n")  # LOOKUP_TABLE
        fields[name] = np.frombuffer(rest[:8 * n], dtype=">f8").astype(np.float64)
        rest = rest[8 * n + 1:]
    return dims, fields


def test_structured_points_roundtrip(tmp_path):
    nx, ny, nz = 5, 4, 3
    a = np.arange(nx * ny * nz, dtype=np.float64).reshape(nx, ny, nz) * 0.5
    path = str(tmp_path / "a.vtk")
    write_structured_points(path, (nx, ny, nz), {"a": a, "b": -a})
    dims, fields = _read(path)
    assert dims == [nx + 1, ny + 1, nz + 1]
    assert np.array_equal(fields["a"].reshape(nz, ny, nx).transpose(2, 1, 0), a)
    assert np.array_equal(fields["b"], -fields["a"])
    with pytest.raises(ValueError):
        write_structured_points(str(tmp_path / "a.vtu"), (nx, ny, nz), {"a": a})


@pytest.mark.gpu
def test_export_vtk_fields(dev, tmp_path):
    import lbm_jax_autodiff_amd as lbm
    nx, ny, nz = 8, 6, 4
    g = torch.Generator().manual_seed(0)
    rho = (1.0 + 0.01 * torch.randn((nx, ny, nz), generator=g, dtype=torch.float64)).to(dev)
    u = (0.02 * torch.randn((3, nx, ny, nz), generator=g, dtype=torch.float64)).to(dev)
    f = lbm.equilibrium(rho, u)
    solid = torch.zeros((nx, ny, nz), dtype=torch.int32, device=dev)
    solid[2:4, 2:4, 1:3] = 1
    path = str(tmp_path / "flow.vtk")
    export_vtk(path, f, solid=solid, porosity=solid.to(torch.float64) * 0.5)
    _, fields = _read(path)
    back = {k: v.reshape(nz, ny, nx).transpose(2, 1, 0) for k, v in fields.items()}
    assert np.allclose(back["density"], rho.cpu().numpy(), atol=1e-12)
    assert np.allclose(back["velocity_x"], u[0].cpu().numpy(), atol=1e-12)
    assert np.allclose(back["velocity_magnitude"], u.pow(2).sum(0).sqrt().cpu().numpy(), atol=1e-12)
    assert np.array_equal(back["solid"], solid.cpu().numpy().astype(np.float64))
    assert set(back) == {"density", "velocity_x", "velocity_y", "velocity_z", "velocity_magnitude", "solid",
                         "porosity"}
