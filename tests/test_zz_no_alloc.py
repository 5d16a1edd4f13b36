"""The no-allocation chained step (lbm_step_chain_forces_noalloc) against the
allocating pair lbm_step_chain + lbm_force_reduce, through the C ABI: the
same kernels in the same order, so states and forces are bit-identical.  Two
calls in a row, the second with a smaller grid and fewer steps (the shared
partials array keeps the first call's contents)."""

import pytest
import torch

import lbm_jax_autodiff_amd as lbm
from lbm_jax_autodiff_amd import _lib as L
from lbm_jax_autodiff_amd.engine import Lattice

pytestmark = pytest.mark.gpu


def _case(dev, nx, ny, nz, storage):
    solid, q = lbm.create_sphere(nx, ny, nz, (0.0, 0.0, 0.0), 0.5, device=dev)
    lat = Lattice(nx, ny, nz, tau=0.56, inlet_velocity=(0.05, 0.0, 0.0), collision="mrt", solid=solid, q=q,
                  guard="pre", guard_velocity=(0.05, 0.0, 0.0), storage=storage, offset="inlet", device=dev)
    u = torch.zeros((3, nx, ny, nz), dtype=torch.float32, device=dev)
    u[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=torch.float32, device=dev), u)
    p0 = lat.new_state()
    lat.collide(lat.to_state(f0), p0)
    return lat, p0


def _reference(lat, p0, n):
    states = [p0] + [lat.new_state() for _ in range(n)]
    fpart = torch.zeros((n, max(lat.slots, 1), 3), dtype=torch.float64, device=lat.device)
    lat.chain(states, fpart if lat.slots else None)
    return states[-1], lat.reduce_forces(fpart, n)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_noalloc_chain_matches_allocating_chain(dev, storage):
    for (nx, ny, nz), n in (((64, 32, 32), 40), ((32, 16, 16), 9)):
        lat, p0 = _case(dev, nx, ny, nz, storage)
        ref_state, ref_force = _reference(lat, p0, n)
        states = [p0] + [lat.new_state() for _ in range(n)]
        force = torch.full((n, 3), float("nan"), dtype=torch.float64, device=dev)
        lat.chain_forces_noalloc(states, force)
        torch.cuda.synchronize()
        assert lat.slots > 0
        assert torch.equal(force, ref_force), (nx, n)
        assert torch.equal(states[-1], ref_state), (nx, n)


def test_noalloc_rejects_too_many_steps(dev):
    lat, p0 = _case(dev, 64, 32, 32, "f32")
    n = L.LBM_NOALLOC_PARTIALS // (3 * lat.slots) + 1
    with pytest.raises(ValueError, match="exceed"):
        lat.chain_forces_noalloc([p0] * (n + 1), torch.zeros((n, 3), dtype=torch.float64, device=dev))
    # the library's own check, below the Python one
    desc = lat.desc
    force = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    rc = L.load().lbm_step_chain_forces_noalloc(desc, n, L.ptr_array([p0] * (n + 1)), L.ptr(force),
                                                L.stream_handle(dev))
    assert rc == 1001
