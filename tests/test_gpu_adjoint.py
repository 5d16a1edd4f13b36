"""Hand-written HIP adjoint against torch autograd of the FP64 oracle, the
reference's published FP64 gradient KATs, finite differences and the exact
zero of the halfway/hard-mask mode."""

import itertools

import pytest
import torch

from lbm_jax_autodiff_amd import adjoint as A
import lbm_jax_autodiff_amd as lbm
from oracle import lbm_oracle as O
from _helpers import near_eq_state, random_q, cpu

pytestmark = pytest.mark.gpu

NX, NY, NZ = 12, 8, 6


def _weights(seed, shape):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(shape, generator=g, dtype=torch.float64)


CASES = list(itertools.product(["bgk", "mrt"], ["none", "hard", "soft"], ["none", "post", "pre"], [False, True]))


def _check_against_oracle(dev, dims, collision, bounce, guard, smag, cells_per_thread=0):
    NX, NY, NZ = dims
    steps = 5
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=4, p_wall=0.1)
    g = torch.Generator().manual_seed(2)
    solid = (torch.rand((NX, NY, NZ), generator=g) < 0.1).to(torch.int32)
    phi = 0.3 * torch.rand((NX, NY, NZ), generator=g, dtype=torch.float64)
    wS, wP, wF = _weights(1, (19, NX, NY, NZ)), _weights(2, (19, NX, NY, NZ)), _weights(3, (steps, 3))
    u0 = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64)

    # oracle with torch autograd
    f0r = f0.clone().requires_grad_(True)
    qr = q.clone().requires_grad_(True)
    phr = phi.clone().requires_grad_(True)
    ur = u0.clone().requires_grad_(True)
    feq_ref = O.equilibrium(torch.ones((NX, NY, NZ), dtype=torch.float64),
                            torch.stack([torch.full((NX, NY, NZ), 0.05, dtype=torch.float64),
                                         torch.zeros((NX, NY, NZ), dtype=torch.float64),
                                         torch.zeros((NX, NY, NZ), dtype=torch.float64)]))
    s_r, p_r, F_r = O.rollout(f0r, steps, forces=True, tau=torch.tensor(0.6, dtype=torch.float64), velocity=ur,
                              q=qr, solid=solid if bounce == "hard" else None,
                              porosity=phr if bounce == "soft" else None, collision=collision,
                              smagorinsky_cs=0.1 if smag else None, guard=guard, f_eq_ref=feq_ref)
    loss_r = (s_r * wS).sum() + (p_r * wP).sum() + (F_r * wF).sum()
    inputs = [f0r, qr, ur] + ([phr] if bounce == "soft" else [])
    grads_r = torch.autograd.grad(loss_r, inputs, allow_unused=True)

    f0d = f0.to(dev).requires_grad_(True)
    qd = q.to(dev).requires_grad_(True)
    phd = phi.to(dev).requires_grad_(True)
    ud = u0.to(dev).requires_grad_(True)
    s, p, F = A.rollout(f0d, steps, tau=0.6, q=qd, porosity=phd if bounce == "soft" else None,
                        solid=solid.to(dev) if bounce == "hard" else None, inlet_velocity=ud,
                        collision=collision, smagorinsky_cs=0.1 if smag else None, guard=guard,
                        guard_velocity=(0.05, 0.0, 0.0), storage="f64", cells_per_thread=cells_per_thread)
    loss = (s * wS.to(dev)).sum() + (p * wP.to(dev)).sum() + (F * wF.to(dev)).sum()
    inputs_d = [f0d, qd, ud] + ([phd] if bounce == "soft" else [])
    grads = torch.autograd.grad(loss, inputs_d, allow_unused=True)
    assert abs(float(loss) - float(loss_r)) < 1e-10 * max(1.0, abs(float(loss_r)))
    errs = {}
    for name, gd, gr in zip(["f0", "q", "u", "phi"], grads, grads_r):
        gr = torch.zeros_like(cpu(gd)) if gr is None else gr
        if name == "q":  # only link entries (q >= 0) carry a gradient
            gr = torch.where(q >= 0, gr, torch.zeros_like(gr))
        err = float((cpu(gd).to(torch.float64) - gr).abs().max())
        scale = float(gr.abs().max()) + 1e-30
        errs[name] = (err, scale, cpu(gd).reshape(-1)[:3].tolist(), gr.reshape(-1)[:3].tolist())
    for name, (err, scale, _, _) in errs.items():
        assert err < 1e-10 * max(1.0, scale), errs


@pytest.mark.parametrize("collision,bounce,guard,smag", CASES)
def test_adjoint_matches_oracle_autograd_f64(dev, collision, bounce, guard, smag):
    _check_against_oracle(dev, (NX, NY, NZ), collision, bounce, guard, smag)


@pytest.mark.parametrize("collision,bounce,guard,smag", CASES)
def test_adjoint_row_sweep_matches_oracle_autograd_f64(dev, collision, bounce, guard, smag):
    """nx = 64: every wave's cells share one row and the dense adjoint sweep
    takes its scalar-addressed path (and reads the wrapped neighbour where
    Zou-He overwrites the pulled value)."""
    _check_against_oracle(dev, (64, 6, 4), collision, bounce, guard, smag)


# The two sweeps are compiled separately and the compiler contracts a * b + c
# into fused multiply-adds where it finds them, so their roundings may differ
# by a few units in the last place; the tolerances allow that and nothing
# more.  Half storage rounds the states to 16 bits, so a last-place
# difference of the float32 arithmetic can flip a stored value.
@pytest.mark.parametrize("collision,bounce,guard", [("mrt", "hard", "pre"), ("bgk", "soft", "post"),
                                                   ("mrt", "soft", "none")])
def test_adjoint_row_sweep_two_waves_per_row_matches_oracle(dev, collision, bounce, guard):
    """nx = 128: two waves per row and two rows per workgroup in the
    scalar-addressed adjoint sweep."""
    _check_against_oracle(dev, (128, 4, 3), collision, bounce, guard, True)


@pytest.mark.parametrize("storage,rtol", [("f64", 1e-13), ("f32", 2e-6), ("f16", 5e-3)])
def test_row_sweep_gradient_matches_per_lane_addressing(dev, storage, rtol):
    """The scalar-addressed sweeps (default where nx % 64 == 0) and the
    per-lane-addressed ones (cells_per_thread=1) give the same states, forces
    and gradients (q, porosity, f0, inlet velocity) to rounding."""
    nx, ny, nz = 64, 20, 12
    dt = torch.float64 if storage == "f64" else torch.float32
    por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-2.0, 0.0, 0.0), 0.5, 10.0, device=dev, out_dtype=dt)
    u = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
    u[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), u)
    out = []
    for cpt in (0, 1):
        qq = q.clone().requires_grad_(True)
        pp = por.clone().requires_grad_(True)
        ff = f0.clone().requires_grad_(True)
        uu = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64, device=dev, requires_grad=True)
        S, P, F = A.rollout(ff, 13, tau=0.55, q=qq, porosity=pp, inlet_velocity=uu, collision="mrt", guard="pre",
                            smagorinsky_cs=0.1, storage=storage, offset="inlet", checkpoint=4, cells_per_thread=cpt)
        loss = F[:, 0].sum() + S.double().sum() * 1e-3
        g = torch.autograd.grad(loss, (qq, pp, ff, uu))
        out.append([cpu(t).double() for t in (S, P, F) + g])
    for name, a, b in zip(["S", "P", "F", "dq", "dphi", "df0", "du"], *out):
        scale = float(a.abs().max())
        assert scale > 0.0, name
        assert float((a - b).abs().max()) <= rtol * scale, name


@pytest.mark.parametrize("guard", ["none", "pre"])
def test_adjoint_is_deterministic(dev, guard):
    """Two identical backward passes give bit-identical gradients (no atomics,
    every cotangent entry written exactly once per step)."""
    steps = 5
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=4, p_wall=0.1)
    out = []
    for rep in range(3):
        # perturb the caching allocator between repetitions
        junk = torch.full((19 * NX * NY * NZ * 4,), float(rep + 1), dtype=torch.float64, device=dev)
        del junk
        f0d = f0.to(dev).requires_grad_(True)
        qd = q.to(dev).requires_grad_(True)
        ud = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64, device=dev, requires_grad=True)
        s, p, F = A.rollout(f0d, steps, tau=0.6, q=qd, inlet_velocity=ud, collision="mrt", guard=guard,
                            guard_velocity=(0.05, 0.0, 0.0), storage="f64")
        loss = s.sum() + (p * p).sum() + F.sum()
        out.append([cpu(g) for g in torch.autograd.grad(loss, [f0d, qd, ud])])
    for rep in (1, 2):
        for a, b in zip(out[0], out[rep]):
            assert torch.equal(a, b)


def test_kat_inlet_velocity_gradient(dev):
    """FP64 d mean(rho[2:-2]) / dU after 5 BGK steps = 0.611238 (paper/paper.tex:502;
    protocol tests/test_gradients_f64.py:95-127)."""
    nx, ny, nz = 8, 4, 4
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=torch.float64, device=dev),
                         torch.zeros((3, nx, ny, nz), dtype=torch.float64, device=dev))
    U = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64, device=dev, requires_grad=True)
    s, _, _ = A.rollout(f0, 5, tau=0.8, q=None, inlet_velocity=U, collision="bgk", storage="f64")
    loss = s[:, 2:-2].sum(0).mean()
    (g,) = torch.autograd.grad(loss, U)
    assert abs(float(g[0]) - 0.611238) < 5e-7


def _kat_cd(r, dev):
    """Cd(r) of tests/test_gradients_f64.py:129-207 on the GPU in FP64."""
    nx, ny, nz = 32, 16, 16
    u_in = 0.05
    D = 32 * 0.5 / 4.0  # tests/test_gradients_f64.py:137
    tau = 3.0 * (u_in * D / 50.0) + 0.5
    por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), r, 20.0, coord_dtype=torch.float64,
                                               out_dtype=torch.float64, disc_floor=1e-20, denom_eps=1e-30,
                                               device=dev)
    uu = torch.zeros((3, nx, ny, nz), dtype=torch.float64, device=dev)
    uu[0] = u_in
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=torch.float64, device=dev), uu)
    _, _, F = A.rollout(f0, 10, tau=tau, q=q, porosity=por, inlet_velocity=(u_in, 0.0, 0.0), collision="bgk",
                        storage="f64")
    area = torch.clamp(lbm.projected_area(solid, 0).to(torch.float64), min=1.0)
    return A.drag_series(F, u_in, area)[-1]


def test_kat_bouzidi_radius_gradient(dev):
    """FP64 Bouzidi dCd/dr over 10 steps = 93.33275 (paper/paper.tex:503)."""
    r = torch.tensor(0.5, dtype=torch.float64, device=dev, requires_grad=True)
    cd = _kat_cd(r, dev)
    (g,) = torch.autograd.grad(cd, r)
    assert abs(float(g) - 93.33275) < 5e-5
    gref, cdref = O.kat_bouzidi_grad()
    assert abs(float(cd) - cdref) < 1e-10
    # central finite difference of the same FP64 kernels
    h = 1e-6
    fd = (float(_kat_cd(torch.tensor(0.5 + h, dtype=torch.float64, device=dev), dev))
          - float(_kat_cd(torch.tensor(0.5 - h, dtype=torch.float64, device=dev), dev))) / (2 * h)
    assert abs(fd - float(g)) / abs(float(g)) < 1e-6


def test_halfway_hard_gradient_is_exactly_zero(dev):
    """q = 1/2 and a hard mask make Cd independent of r: dCd/dr = 0.0 exactly
    (README.md:35, paper/paper.tex:651-662)."""
    nx, ny, nz = 32, 16, 16
    r = torch.tensor(0.5, dtype=torch.float32, device=dev, requires_grad=True)
    por, solid, _ = lbm.smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), r, 20.0, device=dev)
    q = lbm.compute_q_from_solid(solid, nx, ny, nz, 0.5)
    # hard mode: porosity = solid (examples/11_optimisation_v2.py:62-65); the
    # 0 * porosity(r) term keeps r on the graph so the whole adjoint runs
    hard = solid.to(torch.float32) + 0.0 * por
    uu = torch.zeros((3, nx, ny, nz), dtype=torch.float32, device=dev)
    uu[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=torch.float32, device=dev), uu)
    _, _, F = A.rollout(f0, 20, tau=0.56, q=q, porosity=hard, collision="mrt")
    cd = A.drag_series(F, 0.05, torch.clamp(lbm.projected_area(solid), min=1.0))[-1]
    assert float(cd) > 0.0
    (g,) = torch.autograd.grad(cd, r)
    assert float(g) == 0.0


@pytest.mark.parametrize("policy", [3, (7, 3)])
def test_checkpointed_gradient_equals_store_all(dev, policy):
    nx, ny, nz = 24, 12, 12
    solid, q = lbm.create_sphere(nx, ny, nz, (-1.0, 0.0, 0.0), 0.6, device=dev)
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), device=dev), torch.zeros((3, nx, ny, nz), device=dev))
    out = []
    for ck in ("all", policy):
        qq = q.clone().requires_grad_(True)
        _, _, F = A.rollout(f0, 23, tau=0.55, q=qq, solid=solid, collision="mrt", guard="pre", checkpoint=ck)
        (g,) = torch.autograd.grad(F[:, 0].sum(), qq)
        out.append(cpu(g))
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_stochastic_rounding_recompute_reproduces_forward(dev, storage):
    """With stochastic rounding the dither of every stored value is a hash of
    (seed, time index, cell, population): one-level and two-level
    checkpointed rollouts recompute the forward's states bit for bit, so
    forces, states and gradients equal the store-all run's; another seed
    gives other states; the result stays close to the float32 run."""
    nx, ny, nz = 24, 12, 12
    por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-1.0, 0.1, 0.0), 0.6, 10.0, device=dev)
    u = torch.zeros((3, nx, ny, nz), device=dev)
    u[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), device=dev), u)
    out = []
    for ck, seed, st in (("all", 3, storage), (5, 3, storage), ((7, 3), 3, storage), ("all", 4, storage),
                         ("all", 3, "f32")):
        qq = q.clone().requires_grad_(True)
        pp = por.clone().requires_grad_(True)
        S, P, F = A.rollout(f0, 23, tau=0.56, q=qq, porosity=pp, collision="mrt", guard="pre", storage=st,
                            offset="inlet", checkpoint=ck, rounding="stochastic", seed=seed,
                            smagorinsky_cs=0.1)
        loss = F[:, 0].sum() + (S.double() ** 2).sum() * 1e-3
        out.append([cpu(t).double() for t in (S, F) + torch.autograd.grad(loss, (qq, pp))])
    for a, b in zip(out[0], out[1]):
        assert torch.equal(a, b)
    for a, b in zip(out[0], out[2]):
        assert torch.equal(a, b)
    assert not torch.equal(out[0][0], out[3][0])
    tol = 1e-2 if storage == "f16" else 5e-2
    for a, b in zip(out[0], out[4]):
        assert float((a - b).abs().max()) <= tol * max(1.0, float(b.abs().max()))


@pytest.mark.parametrize("storage,tol", [("f64", 1e-13), ("f32", 2e-6), ("f16", 2e-3)])
@pytest.mark.parametrize("checkpoint", ["all", (7, 3)])
def test_folded_link_additions_match_scatter(dev, storage, tol, checkpoint):
    """Folding each adjoint step's Bouzidi and force additions into the next
    step's loads gives the gradients of the separate per-entry scatter (the
    same sums in another order), for f0, q, porosity and the inlet velocity;
    high and low branches, soft and hard bounce, a chain of several steps."""
    nx, ny, nz = 24, 12, 12
    dt = torch.float64 if storage == "f64" else torch.float32
    por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-1.0, 0.1, 0.0), 0.6, 10.0, device=dev, out_dtype=dt)
    u = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
    u[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), u)
    qv = q[q >= 0]
    assert bool((qv < 0.5).any()) and bool((qv >= 0.5).any())
    for bounce in ("soft", "hard"):
        grads = []
        for mode in ("scatter", "fold"):
            ff = f0.clone().requires_grad_(True)
            qq = q.clone().requires_grad_(True)
            pp = por.clone().requires_grad_(True)
            vel = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64, device=dev, requires_grad=True)
            kw = dict(porosity=pp) if bounce == "soft" else dict(solid=solid)
            S, P, F = A.rollout(ff, 17, tau=0.56, q=qq, inlet_velocity=vel, collision="mrt", guard="pre",
                                guard_velocity=(0.05, 0.0, 0.0), storage=storage, offset="inlet",
                                checkpoint=checkpoint, link_additions=mode, **kw)
            loss = F[:, 0].sum() + (S.double() ** 2).sum() * 1e-3
            wrt = [ff, qq, vel] + ([pp] if bounce == "soft" else [])
            grads.append([cpu(g).double() for g in torch.autograd.grad(loss, wrt)])
        for a, b in zip(*grads):
            scale = max(1.0, float(a.abs().max()))
            assert float((a - b).abs().max()) <= tol * scale, (bounce, float((a - b).abs().max()), scale)


@pytest.mark.parametrize("storage", ["f32", "f16", "f64"])
def test_gradient_independent_of_state_layout(dev, storage):
    """Row-interleaved and populations-apart buffers give bit-identical
    forces and gradients (q, porosity, f0)."""
    nx, ny, nz = 24, 12, 12
    dt = torch.float64 if storage == "f64" else torch.float32
    por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-1.0, 0.0, 0.0), 0.6, 10.0, device=dev, out_dtype=dt)
    u = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
    u[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), u)
    out = []
    for layout in ("rows", "apart"):
        qq = q.clone().requires_grad_(True)
        pp = por.clone().requires_grad_(True)
        ff = f0.clone().requires_grad_(True)
        S, P, F = A.rollout(ff, 17, tau=0.55, q=qq, porosity=pp, collision="mrt", guard="pre", storage=storage,
                            offset="inlet", checkpoint=5, layout=layout)
        loss = F[:, 0].sum() + S.double().sum() * 1e-3
        g = torch.autograd.grad(loss, (qq, pp, ff))
        out.append([cpu(t) for t in (S, P, F) + g])
    for a, b in zip(*out):
        assert torch.equal(a, b)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-4), ("f16", 2e-2)])
def test_reduced_precision_gradient_close_to_f64(dev, storage, tol):
    nx, ny, nz = 32, 16, 16
    res = {}
    for st in ("f64", storage):
        dt = torch.float64 if st == "f64" else torch.float32
        r = torch.tensor(0.5, dtype=dt, device=dev, requires_grad=True)
        por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-1.0, 0.0, 0.0), r, 20.0, out_dtype=dt, device=dev)
        hard = solid.to(dt)
        uu = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
        f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), uu)
        _, _, F = A.rollout(f0, 60, tau=0.55, q=q, porosity=hard, collision="mrt", storage=st, offset="inlet")
        cd = A.drag_series(F, 0.05, torch.clamp(lbm.projected_area(solid).to(torch.float64), min=1.0))[-20:].mean()
        (g,) = torch.autograd.grad(cd, r)
        res[st] = (float(cd), float(g))
    assert abs(res[storage][1] - res["f64"][1]) <= tol * abs(res["f64"][1])


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_hard_mask_equals_binary_porosity_blend(dev, storage):
    """The soft blend fed porosity = solid in {0, 1} (examples/11:62-65) and the
    hard-mask reversal give the same forces bit for bit and the same dF/dq."""
    nx, ny, nz = 32, 16, 16
    dt = torch.float32
    r = torch.tensor(0.5, dtype=dt, device=dev)
    _, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-1.0, 0.0, 0.0), r, 20.0, out_dtype=dt, device=dev)
    uu = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
    uu[0] = 0.05
    f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), uu)
    out = []
    for mode in ("blend", "mask"):
        qq = q.detach().clone().requires_grad_(True)
        kw = dict(porosity=solid.to(dt)) if mode == "blend" else dict(solid=solid)
        s, p, F = A.rollout(f0, 30, tau=0.55, q=qq, collision="mrt", guard="pre", storage=storage,
                            offset="inlet", **kw)
        (g,) = torch.autograd.grad(F[-10:, 0].sum(), qq)
        out.append((cpu(s), cpu(F), cpu(g)))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])
    assert float((out[0][2] - out[1][2]).abs().max()) <= 1e-6 * float(out[1][2].abs().max())


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_half_cotangent_storage_close_to_float(dev, storage):
    """Scaled IEEE-half cotangents give the float-cotangent gradient to the
    half rounding level, and the loss is unchanged."""
    nx, ny, nz = 32, 16, 16
    dt = torch.float32
    out = {}
    for lam in ("f32", "f16"):
        r = torch.tensor(0.5, dtype=dt, device=dev, requires_grad=True)
        _, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (-1.0, 0.0, 0.0), r, 20.0, out_dtype=dt, device=dev)
        uu = torch.zeros((3, nx, ny, nz), dtype=dt, device=dev)
        uu[0] = 0.05
        f0 = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=dt, device=dev), uu)
        _, _, F = A.rollout(f0, 80, tau=0.55, q=q, solid=solid, collision="mrt", guard="pre", storage=storage,
                            offset="inlet", checkpoint=16, adjoint_storage=lam)
        cd = A.drag_series(F, 0.05, torch.clamp(lbm.projected_area(solid).to(torch.float64), min=1.0))[-30:].mean()
        (g,) = torch.autograd.grad(cd, r)
        out[lam] = (float(cd), float(g))
    assert out["f16"][0] == out["f32"][0]
    assert abs(out["f16"][1] - out["f32"][1]) <= 2e-3 * abs(out["f32"][1])


def test_half_cotangent_storage_rejected_for_f64(dev):
    f0 = lbm.equilibrium(torch.ones((8, 6, 4), dtype=torch.float64, device=dev),
                         torch.zeros((3, 8, 6, 4), dtype=torch.float64, device=dev))
    with pytest.raises(ValueError):
        A.rollout(f0, 3, tau=0.6, storage="f64", adjoint_storage="f16")
