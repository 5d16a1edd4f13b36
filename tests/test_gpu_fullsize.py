"""Properties that hold at any size, checked at the 256x128x128 benchmark size
(the oracle is too slow there): the fused kernel's per-step force equals the
independent per-operator momentum exchange, state layouts are bit-identical,
float32 tracks float64 (states, forces and the adjoint's dCd/dr),
checkpointed and store-all gradients are bit-identical, and the halfway hard
mask gives an exactly zero dCd/dr."""

import pytest
import torch

import lbm_jax_autodiff_amd as lbm
from lbm_jax_autodiff_amd import adjoint as A
from lbm_jax_autodiff_amd.engine import Lattice

pytestmark = pytest.mark.gpu

NX, NY, NZ = 256, 128, 128
U = 0.05
TAU = 0.548


def _sphere(dev):
    return lbm.create_sphere(NX, NY, NZ, (0.0, 0.0, 0.0), 0.5, device=dev)


def _f0(dev, dt=torch.float32):
    u = torch.zeros((3, NX, NY, NZ), dtype=dt, device=dev)
    u[0] = U
    return lbm.equilibrium(torch.ones((NX, NY, NZ), dtype=dt, device=dev), u)


def _lattice(dev, storage="f32", layout="rows"):
    solid, q = _sphere(dev)
    dt = torch.float64 if storage == "f64" else torch.float32
    return Lattice(NX, NY, NZ, tau=TAU, inlet_velocity=(U, 0.0, 0.0), collision="mrt", solid=solid, q=q.to(dt),
                   guard="pre", guard_velocity=(U, 0.0, 0.0), storage=storage, offset="inlet", device=dev,
                   layout=layout), q.to(dt)


# Both paths add the per-link terms in float64; the terms themselves are computed in
# the storage type with a different association, so float32 agrees to a few float32
# roundings of the largest link term (measured 2.4e-8 of |F| at 20 steps).
@pytest.mark.parametrize("storage,tol", [("f64", 1e-12), ("f32", 1e-6)])
def test_fused_force_equals_per_operator_momentum_exchange(dev, storage, tol):
    lat, q = _lattice(dev, storage)
    dt = torch.float64 if storage == "f64" else torch.float32
    S, P, F = lat.run(_f0(dev, dt), 20)
    ref = lbm.momentum_exchange(P, S, q).to(torch.float64)
    got = F[-1]
    assert float((got - ref).abs().max()) <= tol * float(ref.abs().max()), (got, ref)


def test_state_layouts_bit_identical(dev):
    outs = []
    for layout in ("rows", "apart"):
        lat, _ = _lattice(dev, "f16", layout)
        outs.append(lat.run(_f0(dev), 12))
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_float32_tracks_float64(dev):
    S32, _, F32 = _lattice(dev, "f32")[0].run(_f0(dev), 200)
    S64, _, F64 = _lattice(dev, "f64")[0].run(_f0(dev, torch.float64), 200)
    assert float((S32.double() - S64).abs().max()) < 2e-6
    assert float((F32 - F64).abs().max()) < 1e-6 * float(F64.abs().max()) + 1e-9


def _grad(dev, steps, checkpoint, halfway=False, storage="f32"):
    dt = torch.float64 if storage == "f64" else torch.float32
    r = torch.tensor(0.5, dtype=torch.float32, device=dev, requires_grad=True)
    _, solid, q = lbm.smooth_sphere_geometry(NX, NY, NZ, (0.0, 0.0, 0.0), r, 20.0, device=dev, out_dtype=dt)
    porosity = None
    if halfway:
        # q = 1/2 on every link and the reference's 0 * porosity(r) term
        por, _, _ = lbm.smooth_sphere_geometry(NX, NY, NZ, (0.0, 0.0, 0.0), r, 20.0, device=dev)
        q = lbm.compute_q_from_solid(solid, NX, NY, NZ, 0.5)
        porosity, solid = solid.to(torch.float32) + 0.0 * por, None
    _, _, F = A.rollout(_f0(dev, dt), steps, tau=TAU, q=q, solid=solid, porosity=porosity,
                        inlet_velocity=(U, 0.0, 0.0), collision="mrt", guard="pre", guard_velocity=(U, 0.0, 0.0),
                        offset="inlet", checkpoint=checkpoint, storage=storage)
    area = torch.clamp(lbm.projected_area(lbm.create_sphere(NX, NY, NZ, (0.0, 0.0, 0.0), 0.5, device=dev)[0])
                       .to(torch.float64), min=1.0)
    loss = A.drag_series(F, U, area)[-max(1, steps // 3):].mean()
    (g,) = torch.autograd.grad(loss, r)
    return g


def test_checkpointed_gradient_bit_identical_to_store_all(dev):
    g_all = _grad(dev, 45, "all")
    assert torch.equal(g_all, _grad(dev, 45, 10))
    assert torch.equal(g_all, _grad(dev, 45, (20, 6)))


def test_halfway_hard_gradient_exactly_zero(dev):
    assert float(_grad(dev, 40, "all", halfway=True)) == 0.0


def test_float32_gradient_matches_float64(dev):
    """The float32 adjoint sweep of grids of 2^21 cells or more runs in
    512-thread workgroups, the float64 sweep in 256-thread ones: an
    independent check of the 512-thread launch (40 steps, store-all)."""
    g32 = float(_grad(dev, 40, "all"))
    g64 = float(_grad(dev, 40, "all", storage="f64"))
    print(f"256x128x128 40 steps: dCd/dr f32 {g32:.10f} f64 {g64:.10f} rel {abs(g32 / g64 - 1.0):.2e}")
    assert abs(g32 / g64 - 1.0) < 1e-4
