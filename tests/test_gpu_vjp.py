"""Reverse mode of the per-operator surface against torch autograd of the
oracle (float64), operator by operator and through reference-style
compositions, including the reference's published gradient values computed
through the per-operator path."""

import json
import os

import pytest
import torch

import lbm_jax_autodiff_amd as lbm
from oracle import lbm_oracle as O
from _helpers import near_eq_state, random_q, cpu

pytestmark = pytest.mark.gpu

NX, NY, NZ = 10, 6, 5
DT = torch.float64

with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kats.json")) as _fh:
    KAT = json.load(_fh)


def _grads(fn, inputs, seed=11):
    """(outputs, d<w, fn(inputs)>/d inputs) for a seeded random cotangent w."""
    out = fn(*inputs)
    outs = out if isinstance(out, tuple) else (out,)
    g = torch.Generator().manual_seed(seed)
    loss = 0.0
    for o in outs:
        w = torch.randn(o.shape, generator=g, dtype=torch.float64).to(o.device, o.dtype)
        loss = loss + (o * w).sum()
    want = [t for t in inputs if isinstance(t, torch.Tensor) and t.requires_grad]
    return outs, torch.autograd.grad(loss, want)


def _pair(dev, *tensors):
    """CPU leaf copies and device leaf copies (requires_grad where asked)."""
    cpu_in, dev_in = [], []
    for t, rg in tensors:
        if isinstance(t, torch.Tensor):
            a = t.detach().clone().requires_grad_(rg)
            b = t.detach().clone().to(dev).requires_grad_(rg)
        else:
            a = b = t
        cpu_in.append(a)
        dev_in.append(b)
    return cpu_in, dev_in


def _check(dev, ref_fn, gpu_fn, tensors, tol=1e-12, seed=11):
    ci, di = _pair(dev, *tensors)
    ro, rg = _grads(ref_fn, ci, seed)
    go, gg = _grads(gpu_fn, di, seed)
    for a, b in zip(ro, go):
        assert float((cpu(b).double() - a.detach().double()).abs().max()) <= tol * max(1.0, float(a.abs().max()))
    assert len(rg) == len(gg)
    for a, b in zip(rg, gg):
        err = float((cpu(b).double() - a.double()).abs().max())
        assert err <= tol * max(1.0, float(a.abs().max())), err


def test_equilibrium_density_velocity_vjp(dev):
    g = torch.Generator().manual_seed(0)
    rho = 1.0 + 0.05 * torch.randn((NX, NY, NZ), generator=g, dtype=DT)
    u = 0.04 * torch.randn((3, NX, NY, NZ), generator=g, dtype=DT)
    _check(dev, O.equilibrium, lbm.equilibrium, [(rho, True), (u, True)])
    f = near_eq_state(NX, NY, NZ)
    _check(dev, O.compute_density, lbm.compute_density, [(f, True)])
    _check(dev, O.compute_velocity, lbm.compute_velocity, [(f, True), (O.compute_density(f), True)])


@pytest.mark.parametrize("op", ["bgk", "mrt"])
@pytest.mark.parametrize("tau_kind", ["number", "scalar_tensor", "field"])
def test_collision_vjp(dev, op, tau_kind):
    f = near_eq_state(NX, NY, NZ, amp=3e-3)
    g = torch.Generator().manual_seed(1)
    if tau_kind == "number":
        tau = (0.61, False)
    elif tau_kind == "scalar_tensor":
        tau = (torch.tensor(0.61, dtype=DT), True)
    else:
        tau = (0.55 + 0.1 * torch.rand((NX, NY, NZ), generator=g, dtype=DT), True)
    _check(dev, getattr(O, op), getattr(lbm, op), [(f, True), tau])


def test_smagorinsky_vjp(dev):
    f = near_eq_state(NX, NY, NZ, amp=3e-3)
    _check(dev, lambda a, t: O.smagorinsky_tau(a, t, 0.1), lambda a, t: lbm.smagorinsky_tau(a, t, 0.1),
           [(f, True), (torch.tensor(0.55, dtype=DT), True)])


@pytest.mark.parametrize("periodic", [True, False])
def test_stream_vjp(dev, periodic):
    f = near_eq_state(NX, NY, NZ, amp=3e-3)
    q = random_q(NX, NY, NZ, seed=3, p_wall=0.25)
    _check(dev, lambda a, b: O.stream(a, b, periodic), lambda a, b: lbm.stream(a, b, periodic),
           [(f, True), (q, True)])


@pytest.mark.parametrize("rho_target", [None, 1.02])
def test_zou_he_vjp(dev, rho_target):
    f = near_eq_state(NX, NY, NZ, amp=3e-3)
    vel = torch.tensor([0.05, 0.01, -0.02], dtype=DT)
    rt = (torch.tensor(rho_target, dtype=DT), True) if rho_target is not None else (None, False)
    _check(dev, O.zou_he_inlet, lbm.zou_he_inlet, [(f, True), (vel, True), rt])
    _check(dev, O.zou_he_outlet, lbm.zou_he_outlet, [(f, True), (torch.tensor(1.01, dtype=DT), True)])


def test_bounce_vjp(dev):
    f = near_eq_state(NX, NY, NZ, amp=3e-3)
    g = torch.Generator().manual_seed(4)
    phi = torch.rand((NX, NY, NZ), generator=g, dtype=DT)
    _check(dev, O.soft_bounce_back, lbm.soft_bounce_back, [(f, True), (phi, True)])
    solid = (torch.rand((NX, NY, NZ), generator=g) < 0.3).to(torch.int32)

    def ref_hard(a, s):
        return torch.where((s >= 1).unsqueeze(0), a[list(O.OPPOSITE)], a)
    _check(dev, ref_hard, lambda a, s: lbm.core.hard_bounce_back(a, s), [(f, True), (solid, False)])


def test_momentum_exchange_and_drag_vjp(dev):
    fp = near_eq_state(NX, NY, NZ, amp=3e-3, seed=5)
    fs = near_eq_state(NX, NY, NZ, amp=3e-3, seed=6)
    q = random_q(NX, NY, NZ, seed=7, p_wall=0.2)
    _check(dev, lambda a, b, c: O.momentum_exchange(a, b, c), lambda a, b, c: lbm.momentum_exchange(a, b, c),
           [(fp, True), (fs, True), (q, False)])

    def ref_dec(a, b, c):
        tot, pr = O.momentum_exchange_decomposed(a, b, c)
        return tot, pr
    _check(dev, ref_dec, lambda a, b, c: lbm.momentum_exchange_decomposed(a, b, c),
           [(fp, True), (fs, True), (q, False)], tol=1e-10)
    force = torch.tensor([[0.01, -0.002, 0.003], [-0.02, 0.0, 0.001]], dtype=DT)
    # a (T, 3) history: the oracle's force[0] is the x row of the transposed history
    _check(dev, lambda F_: O.drag_coefficient(F_.T, 0.05, 12.0), lambda F_: lbm.drag_coefficient(F_, 0.05, 12.0),
           [(force, True)])


def _scan_body(ops, f, tau, q, vel):
    """One step as a reference user writes it (examples/11 scan body):
    collide, stream, inlet, outlet."""
    fc = ops.mrt(f, tau)
    fs = ops.stream(fc, q, True)
    fs = ops.zou_he_inlet(fs, vel)
    return ops.zou_he_outlet(fs, 1.0), fc


def test_composed_rollout_vjp(dev):
    """Six reference-style steps with a force loss: gradients with respect to
    the initial state, q and the inlet velocity."""
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=8, p_wall=0.15)
    vel = torch.tensor([0.05, 0.0, 0.0], dtype=DT)

    def run(ops, mex, f, qq, v):
        fp = f
        for _ in range(6):
            f, fp = _scan_body(ops, f, 0.56, qq, v)
        return mex(fp, f, qq)

    _check(dev, lambda a, b, c: run(O, O.momentum_exchange, a, b, c),
           lambda a, b, c: run(lbm, lbm.momentum_exchange, a, b, c), [(f0, True), (q, True), (vel, True)],
           tol=1e-11)


def test_kat_inlet_velocity_gradient_per_op(dev):
    """FP64 d mean(rho)/dU over 5 BGK steps = 0.611238 (paper/paper.tex:502),
    through the per-operator path (tests/test_gradients_f64.py:95-127)."""
    nx, ny, nz = 8, 4, 4
    U = torch.tensor(0.05, dtype=DT, device=dev, requires_grad=True)
    f = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=DT, device=dev), torch.zeros((3, nx, ny, nz), dtype=DT,
                                                                                      device=dev))
    q = -torch.ones((19, nx, ny, nz), dtype=DT, device=dev)
    vel = torch.stack([U, torch.zeros((), dtype=DT, device=dev), torch.zeros((), dtype=DT, device=dev)])
    for _ in range(5):
        f = lbm.zou_he_outlet(lbm.zou_he_inlet(lbm.stream(lbm.bgk(f, 0.8), q, True), vel), 1.0)
    loss = lbm.compute_density(f)[2:-2].mean()
    (g,) = torch.autograd.grad(loss, U)
    assert abs(float(g) - KAT["d_mean_rho_d_u_inlet"]["value"]) < 5e-7, float(g)


def test_kat_bouzidi_radius_gradient_per_op(dev):
    """FP64 Bouzidi dCd/dr over 10 steps = 93.33275 (paper/paper.tex:503),
    through the per-operator path: soft porosity and ray q(r) from
    smooth_sphere_geometry (tests/test_gradients_f64.py:129-217)."""
    nx, ny, nz = 32, 16, 16
    u_in = 0.05
    tau = 3.0 * (u_in * (32 * 0.5 / 4.0) / 50.0) + 0.5
    r = torch.tensor(0.5, dtype=DT, device=dev, requires_grad=True)
    por, solid, q = lbm.smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), r, 20.0, coord_dtype=DT, out_dtype=DT,
                                               disc_floor=1e-20, denom_eps=1e-30)
    u = torch.zeros((3, nx, ny, nz), dtype=DT, device=dev)
    u[0] = u_in
    f = lbm.equilibrium(torch.ones((nx, ny, nz), dtype=DT, device=dev), u)
    fp = f
    for _ in range(10):
        fp = lbm.soft_bounce_back(lbm.bgk(f, tau), por)
        f = lbm.zou_he_outlet(lbm.zou_he_inlet(lbm.stream(fp, q, True), (u_in, 0.0, 0.0)), 1.0)
    force = lbm.momentum_exchange(fp, f, q)
    area = torch.clamp(lbm.projected_area(solid).to(DT), min=1.0)
    cd = lbm.drag_coefficient(force, u_in, area)
    (g,) = torch.autograd.grad(cd, r)
    assert abs(float(g) - KAT["d_cd_d_r_bouzidi_soft"]["value"]) < 5e-5, float(g)


@pytest.mark.parametrize("checkpoint", [True, False])
@pytest.mark.parametrize("collision_mode", [0, 1])
def test_simulate_reverse_mode(dev, checkpoint, collision_mode):
    """simulate() differentiates through the fused rollout: gradients with
    respect to f0, q and the inlet velocity match torch autograd of the
    oracle's simulate (state.py:174-230, post-step guard, hard bounce)."""
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=9, p_wall=0.12)
    g = torch.Generator().manual_seed(10)
    solid = (torch.rand((NX, NY, NZ), generator=g) < 0.08).to(torch.int32)
    vel = torch.tensor([0.05, 0.0, 0.0], dtype=DT)
    n = 7

    def ref(f, qq, v):
        st = O.simulate(O.LBMState(f, solid, qq), O.LBMParams(torch.tensor(0.57, dtype=DT), v, collision_mode),
                        n)
        return st.f

    def gpu(f, qq, v):
        st = lbm.simulate(lbm.LBMState(f, solid.to(dev), qq), lbm.LBMParams(0.57, v, collision_mode), n,
                          checkpoint=checkpoint)
        return st.f

    _check(dev, ref, gpu, [(f0, True), (q, True), (vel, True)], tol=1e-11)


@pytest.mark.parametrize("case", ["clamped", "tau", "tau_clamped_smag", "tau_field"])
def test_simulate_reverse_mode_per_operator_path(dev, case):
    """simulate() with clamped y/z (streaming.py:134-142) or a tau that
    requires grad or varies per cell runs the per-operator composition;
    gradients in f0, q, the inlet velocity and tau (LBMParams leaves that
    jax.grad reaches, state.py:57-84) match torch autograd of the oracle."""
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3)
    q = random_q(NX, NY, NZ, seed=9, p_wall=0.12)
    g = torch.Generator().manual_seed(10)
    solid = (torch.rand((NX, NY, NZ), generator=g) < 0.08).to(torch.int32)
    vel = torch.tensor([0.05, 0.01, 0.0], dtype=DT)
    periodic = case not in ("clamped", "tau_clamped_smag")
    smag = int(case == "tau_clamped_smag")
    coll = 1
    if case == "tau_field":
        tau = 0.55 + 0.2 * torch.rand((NX, NY, NZ), generator=g, dtype=DT)
    else:
        tau = torch.tensor(0.57, dtype=DT)
    tau_grad = case != "clamped"
    n = 6

    def ref(f, qq, v, t):
        st = O.simulate(O.LBMState(f, solid, qq), O.LBMParams(t, v, coll, smag, 0.1, periodic), n)
        return st.f

    def gpu(f, qq, v, t):
        st = lbm.simulate(lbm.LBMState(f, solid.to(dev), qq), lbm.LBMParams(t, v, coll, smag, 0.1, periodic), n,
                          checkpoint=True)
        return st.f

    _check(dev, ref, gpu, [(f0, True), (q, True), (vel, True), (tau, tau_grad)], tol=1e-11)


def test_inlet_velocity_gradient_on_host_tensor(dev):
    """A host (CPU) inlet velocity with requires_grad gets its cotangent on
    the host, through the fused rollout and through the per-operator path."""
    f0 = near_eq_state(NX, NY, NZ, amp=2e-3).to(dev)
    q = random_q(NX, NY, NZ, seed=9, p_wall=0.12).to(dev)
    solid = torch.zeros((NX, NY, NZ), dtype=torch.int32, device=dev)
    for periodic in (True, False):
        vel = torch.tensor([0.05, 0.0, 0.0], dtype=DT, requires_grad=True)
        st = lbm.simulate(lbm.LBMState(f0, solid, q), lbm.LBMParams(0.6, vel, 1, 0, 0.1, periodic), 3)
        (gv,) = torch.autograd.grad(lbm.compute_density(st.f).sum(), vel)
        assert gv.device.type == "cpu" and torch.isfinite(gv).all()


def _geometry_grads(geom, nx, ny, nz, center, radius, dev, weights_seed=12, **kw):
    """d/d(r, c) of <w_p, porosity> + <w_q, q> (and the outputs) for a geometry function."""
    c = torch.tensor(center, dtype=DT, device=dev, requires_grad=True)
    r = torch.tensor(radius, dtype=DT, device=dev, requires_grad=True)
    out = geom(nx, ny, nz, c, r, **kw)
    por, q = (out[0], out[2]) if len(out) == 3 else (None, out[1])
    g = torch.Generator().manual_seed(weights_seed)
    loss = (q * torch.randn(q.shape, generator=g, dtype=DT).to(q.device, q.dtype)).sum()
    if por is not None:
        loss = loss + (por * torch.randn(por.shape, generator=g, dtype=DT).to(por.device, por.dtype)).sum()
    gr, gc = torch.autograd.grad(loss, (r, c))
    return float(gr), [float(v) for v in gc]


@pytest.mark.parametrize("center", [(-0.3, 0.1, -0.05), (0.0, 0.0, 0.0)])
def test_smooth_sphere_center_and_radius_gradients(dev, center):
    """dq/dr, dq/dc, dporosity/dr, dporosity/dc (smooth.py:158-248) against
    torch autograd of the oracle, float64 coordinates."""
    nx, ny, nz = 32, 16, 16

    def ref(nx_, ny_, nz_, c, r):
        return O.smooth_sphere_geometry(nx_, ny_, nz_, c, r, 10.0, coord_dtype=DT)

    def gpu(nx_, ny_, nz_, c, r):
        return lbm.smooth_sphere_geometry(nx_, ny_, nz_, c, r, 10.0, coord_dtype=DT, out_dtype=DT)

    gr_ref, gc_ref = _geometry_grads(ref, nx, ny, nz, center, 0.55, torch.device("cpu"))
    gr, gc = _geometry_grads(gpu, nx, ny, nz, center, 0.55, dev)
    assert abs(gr - gr_ref) <= 1e-10 * max(1.0, abs(gr_ref)), (gr, gr_ref)
    for a, b in zip(gc, gc_ref):
        assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (gc, gc_ref)


@pytest.mark.parametrize("center,radius", [((-0.3, 0.1, -0.05), 0.55), ((-3.7, 1.6, -1.8), 0.5),
                                           ((4.2, 0.0, 0.0), 0.6), ((0.0, -1.9, 1.95), 1.1)])
def test_sphere_q_gradient_sweeps_only_the_sphere_box(dev, center, radius):
    """With no porosity cotangent the geometry's reverse product sweeps only
    the sphere's box of cells; a zero porosity cotangent makes it sweep the
    whole grid.  Both give dq/d(r, c) of the oracle, for spheres inside the
    domain, against and across its faces (clamped boxes)."""
    nx, ny, nz = 32, 16, 16
    g = torch.Generator().manual_seed(14)
    w = torch.randn((19, nx, ny, nz), generator=g, dtype=DT)

    def grads(geom, device, with_por):
        c = torch.tensor(center, dtype=DT, device=device, requires_grad=True)
        r = torch.tensor(radius, dtype=DT, device=device, requires_grad=True)
        por, _, q = geom(c, r)
        loss = (q * w.to(q.device)).sum()
        if with_por:
            loss = loss + 0.0 * por.sum()
        gr, gc = torch.autograd.grad(loss, (r, c))
        return [float(gr)] + [float(v) for v in gc]

    ref = grads(lambda c, r: O.smooth_sphere_geometry(nx, ny, nz, c, r, 10.0, coord_dtype=DT), "cpu", False)
    box = grads(lambda c, r: lbm.smooth_sphere_geometry(nx, ny, nz, c, r, 10.0, coord_dtype=DT, out_dtype=DT),
                dev, False)
    full = grads(lambda c, r: lbm.smooth_sphere_geometry(nx, ny, nz, c, r, 10.0, coord_dtype=DT, out_dtype=DT),
                 dev, True)
    scale = max(1.0, max(abs(v) for v in ref))
    assert max(abs(a - b) for a, b in zip(box, ref)) <= 1e-10 * scale, (box, ref)
    assert max(abs(a - b) for a, b in zip(box, full)) <= 1e-12 * scale, (box, full)


def test_create_sphere_q_gradients(dev):
    """The hard sphere's ray q (primitives.py:56-174) in radius and centre,
    float32 coordinates as in the reference, against oracle autograd."""
    nx, ny, nz = 32, 16, 16
    c0, r0 = (-0.3, 0.1, -0.05), 0.55

    def grads(geom, device, dt):
        c = torch.tensor(c0, dtype=dt, device=device, requires_grad=True)
        r = torch.tensor(r0, dtype=dt, device=device, requires_grad=True)
        solid, q = geom(c, r)
        g = torch.Generator().manual_seed(13)
        w = torch.randn(q.shape, generator=g, dtype=torch.float64).to(q.device, q.dtype)
        return torch.autograd.grad((q * w).sum(), (r, c)), solid

    (gr_ref, gc_ref), s_ref = grads(lambda c, r: O.create_sphere(nx, ny, nz, c, r), "cpu", torch.float32)
    (gr, gc), s = grads(lambda c, r: lbm.create_sphere(nx, ny, nz, c, r, device=dev), dev, torch.float32)
    assert torch.equal(cpu(s), s_ref)
    assert abs(float(gr) - float(gr_ref)) <= 1e-4 * abs(float(gr_ref))
    assert float((cpu(gc) - gc_ref).abs().max()) <= 1e-4 * float(gc_ref.abs().max())
