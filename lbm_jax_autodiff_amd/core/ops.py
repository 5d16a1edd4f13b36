"""jax_lbm.core operators backed by liblbm_hip.so.

Every function keeps the name, argument order, defaults and return shape of
the reference function it replaces (cited per function) and takes / returns
torch tensors on a HIP device.  Outputs are views of lattice storage (x
fastest), indexable exactly like the reference's arrays.  Inputs are never
modified.
"""

from __future__ import annotations

import torch

from .. import _field as F
from .. import _lib as L


def _launch(name, *args):
    L.call(name, *args)


def _compute_density_fwd(f: torch.Tensor) -> torch.Tensor:
    """rho = sum_i f_i.  jax_lbm/core/equilibrium.py:22-48."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    out = F.new_scalar(nx, ny, nz, fs.dtype, fs.device)
    _launch("lbm_density", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(out), nx, ny, nz, L.stream_handle(fs.device))
    return F.scalar_view(out)


def _compute_velocity_fwd(f: torch.Tensor, rho: torch.Tensor) -> torch.Tensor:
    """u = sum_i f_i e_i / rho.  jax_lbm/core/equilibrium.py:51-87."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    rs = F.as_field(rho, fs, (nx, ny, nz))
    out = torch.empty((3, nz, ny, nx), dtype=fs.dtype, device=fs.device)
    _launch("lbm_velocity", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(rs), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.vec_view(out)


def _equilibrium_fwd(rho, u: torch.Tensor) -> torch.Tensor:
    """Second-order D3Q19 equilibrium.  jax_lbm/core/equilibrium.py:90-151."""
    F.require_cuda(u, "u")
    if u.dim() != 4 or u.shape[0] != 3:
        raise ValueError(f"u must have shape (3, nx, ny, nz), got {tuple(u.shape)}")
    nx, ny, nz = int(u.shape[1]), int(u.shape[2]), int(u.shape[3])
    dt = u.dtype if not isinstance(rho, torch.Tensor) else torch.promote_types(u.dtype, rho.dtype)
    us = F.vec_storage(u.to(dt))
    rs = F.as_field(rho, us, (nx, ny, nz))
    out = F.new_pop(nx, ny, nz, dt, us.device)
    _launch("lbm_equilibrium", F.dtype_tag(dt), L.ptr(rs), L.ptr(us), L.ptr(out), nx, ny, nz,
            L.stream_handle(us.device))
    return F.pop_view(out)


def _collide(name, f, tau):
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    out = torch.empty_like(fs)
    tau_field = None
    tau_s = 0.0
    if isinstance(tau, torch.Tensor) and tau.dim() > 0:
        tau_field = F.as_field(tau, fs, (nx, ny, nz))
    else:
        tau_s = F.scalar_float(tau)
    _launch(name, F.dtype_tag(fs.dtype), L.ptr(fs), tau_s, L.ptr(tau_field), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.pop_view(out)


def _bgk_fwd(f: torch.Tensor, tau) -> torch.Tensor:
    """BGK collision; tau scalar or (nx, ny, nz).  jax_lbm/core/collision.py:46-100."""
    return _collide("lbm_bgk", f, tau)


def _mrt_fwd(f: torch.Tensor, tau) -> torch.Tensor:
    """MRT collision (d'Humieres).  jax_lbm/core/collision.py:103-184."""
    return _collide("lbm_mrt", f, tau)


def _smagorinsky_tau_fwd(f: torch.Tensor, tau, cs: float = 0.1) -> torch.Tensor:
    """Effective relaxation time, shape (nx, ny, nz).  jax_lbm/core/collision.py:187-270."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    out = F.new_scalar(nx, ny, nz, fs.dtype, fs.device)
    _launch("lbm_smagorinsky_tau", F.dtype_tag(fs.dtype), L.ptr(fs), F.scalar_float(tau), float(cs), L.ptr(out),
            nx, ny, nz, L.stream_handle(fs.device))
    return F.scalar_view(out)


def _stream_fwd(f: torch.Tensor, q: torch.Tensor, periodic_yz: bool = True) -> torch.Tensor:
    """Pull streaming with Bouzidi links.  jax_lbm/core/streaming.py:40-185."""
    F.require_cuda(f, "f")
    F.require_cuda(q, "q")
    nx, ny, nz = F.dims_of_pop(f)
    F.require_shape(q, "q", f.shape, f.device)
    fs = F.pop_storage(f)
    qs = F.pop_storage(q.to(fs.dtype))
    out = torch.empty_like(fs)
    _launch("lbm_stream", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(qs), int(bool(periodic_yz)), L.ptr(out),
            nx, ny, nz, L.stream_handle(fs.device))
    return F.pop_view(out)


def _vec3(v):
    if isinstance(v, torch.Tensor):
        vals = v.detach().reshape(-1).to("cpu", torch.float64).tolist()
    else:
        vals = [float(x) for x in v]
    if len(vals) != 3:
        raise ValueError("velocity must have 3 components")
    return vals


def _zou_he_inlet_fwd(f: torch.Tensor, velocity, rho_target=None) -> torch.Tensor:
    """Zou-He velocity inlet at x = 0.  jax_lbm/core/boundary.py:35-124."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    out = F.pop_storage(f).clone()
    ux, uy, uz = _vec3(velocity)
    has_rt = rho_target is not None
    rt = F.scalar_float(rho_target) if has_rt else 0.0
    _launch("lbm_zou_he_inlet", F.dtype_tag(out.dtype), L.ptr(out), ux, uy, uz, int(has_rt), rt, nx, ny, nz,
            L.stream_handle(out.device))
    return F.pop_view(out)


def _zou_he_outlet_fwd(f: torch.Tensor, rho_out: float = 1.0) -> torch.Tensor:
    """Zou-He pressure outlet at x = nx-1.  jax_lbm/core/boundary.py:127-200."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    out = F.pop_storage(f).clone()
    _launch("lbm_zou_he_outlet", F.dtype_tag(out.dtype), L.ptr(out), F.scalar_float(rho_out), nx, ny, nz,
            L.stream_handle(out.device))
    return F.pop_view(out)


def _mex(f_post, f_streamed, q, want_pressure):
    F.require_cuda(f_post, "f_post")
    F.require_cuda(f_streamed, "f_streamed")
    F.require_cuda(q, "q")
    nx, ny, nz = F.dims_of_pop(f_post)
    F.require_shape(f_streamed, "f_streamed", f_post.shape, f_post.device)
    F.require_shape(q, "q", f_post.shape, f_post.device)
    ps = F.pop_storage(f_post)
    ss = F.pop_storage(f_streamed.to(ps.dtype))
    qs = F.pop_storage(q.to(ps.dtype))
    ws = torch.empty(int(L.load().lbm_reduce_workspace_bytes(nx, ny, nz)) // 8, dtype=torch.float64,
                     device=ps.device)
    out = torch.empty(6, dtype=torch.float64, device=ps.device)
    _launch("lbm_momentum_exchange", F.dtype_tag(ps.dtype), L.ptr(ps), L.ptr(ss), L.ptr(qs), int(want_pressure),
            L.ptr(out), L.ptr(ws), nx, ny, nz, L.stream_handle(ps.device))
    return out, ps.dtype


def _momentum_exchange_fwd(f_post: torch.Tensor, f_streamed: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """Force on the walls, (Fx, Fy, Fz), accumulated in float64.  jax_lbm/core/forces.py:37-120."""
    out, dt = _mex(f_post, f_streamed, q, False)
    return out[:3].to(dt)


def _momentum_exchange_decomposed_fwd(f_post, f_streamed, q):
    """(total, pressure) forces.  jax_lbm/core/forces.py:123-196."""
    out, dt = _mex(f_post, f_streamed, q, True)
    return out[:3].to(dt), out[3:].to(dt)


def projected_area(solid: torch.Tensor, axis: int = 0) -> torch.Tensor:
    """Count of columns along ``axis`` with a body cell (solid == 1).  jax_lbm/core/forces.py:199-234."""
    F.require_cuda(solid, "solid")
    if solid.dim() != 3:
        raise ValueError("solid must have shape (nx, ny, nz)")
    nx, ny, nz = (int(s) for s in solid.shape)
    ss = F.scalar_storage(solid.to(torch.int32))
    out = torch.empty((), dtype=torch.float32, device=ss.device)
    counter = torch.empty(1, dtype=torch.int32, device=ss.device)
    _launch("lbm_projected_area", L.ptr(ss), int(axis), L.ptr(out), L.ptr(counter), nx, ny, nz,
            L.stream_handle(ss.device))
    return out


def _drag_coefficient_fwd(force: torch.Tensor, inlet_velocity: float, area, rho: float = 1.0) -> torch.Tensor:
    """Cd = |F_x| / (0.5 rho U^2 A + 1e-10).  jax_lbm/core/forces.py:237-276.

    ``force`` may be (3,) or a (T, 3) history; the result is () or (T,).
    """
    F.require_cuda(force, "force")
    fd = force.detach().to(torch.float64).reshape(-1, 3).contiguous()
    if isinstance(area, torch.Tensor):
        a = area.detach().to(device=fd.device, dtype=torch.float64).reshape(1).contiguous()
    else:
        a = torch.full((1,), float(area), dtype=torch.float64, device=fd.device)
    out = torch.empty(fd.shape[0], dtype=torch.float64, device=fd.device)
    _launch("lbm_drag_coefficient", L.ptr(fd), int(fd.shape[0]), F.scalar_float(inlet_velocity), L.ptr(a),
            float(rho), L.ptr(out), L.stream_handle(fd.device))
    out = out.to(force.dtype if force.dtype in (torch.float32, torch.float64) else torch.float64)
    return out.reshape(()) if force.dim() == 1 else out


def _soft_bounce_back_fwd(f_collided: torch.Tensor, porosity: torch.Tensor) -> torch.Tensor:
    """(1 - phi) f + phi f[opp].  jax_lbm/geometry/smooth.py:118-155."""
    F.require_cuda(f_collided, "f_collided")
    nx, ny, nz = F.dims_of_pop(f_collided)
    fs = F.pop_storage(f_collided)
    ps = F.as_field(porosity, fs, (nx, ny, nz))
    out = torch.empty_like(fs)
    _launch("lbm_soft_bounce_back", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(ps), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.pop_view(out)


def _hard_bounce_back_fwd(f_collided: torch.Tensor, solid: torch.Tensor) -> torch.Tensor:
    """f[opp] on cells with solid >= 1 (jax_lbm/state.py:146-149)."""
    F.require_cuda(f_collided, "f_collided")
    nx, ny, nz = F.dims_of_pop(f_collided)
    F.require_shape(solid, "solid", (nx, ny, nz), f_collided.device)
    fs = F.pop_storage(f_collided)
    ss = F.scalar_storage(solid.to(torch.int32))
    out = torch.empty_like(fs)
    _launch("lbm_hard_bounce", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(ss), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.pop_view(out)


def _stability_guard_fwd(f: torch.Tensor) -> torch.Tensor:
    """Reset unstable cells to f_eq(1, 0) (jax_lbm/state.py:158-169)."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    out = torch.empty_like(fs)
    _launch("lbm_stability_guard", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.pop_view(out)


# ---------------------------------------------------------------------------
# Reverse mode.  Each public operator is a torch.autograd.Function whose
# backward calls the operator's hand-written VJP kernel (lbm_*_vjp), so
# compositions of operators differentiate on the device the way jax.grad
# differentiates compositions of the reference's functions.
# ---------------------------------------------------------------------------

def _needs(t) -> bool:
    return isinstance(t, torch.Tensor) and t.requires_grad


def _field_grad(g_storage: torch.Tensor, like):
    """Cotangent of a scalar-or-field input from its per-cell (nz, ny, nx)
    cotangent: the field itself, its sum for a 0-d tensor, None for a number."""
    if not isinstance(like, torch.Tensor):
        return None
    if like.dim() == 3:
        return F.scalar_view(g_storage).to(like.dtype)
    return g_storage.sum().reshape(like.shape).to(like.device, like.dtype)


def _pop_grad(g: torch.Tensor, dtype) -> torch.Tensor:
    return F.pop_storage(g.to(dtype))


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f):
        ctx.dims = F.dims_of_pop(f)
        return _compute_density_fwd(f)

    @staticmethod
    def backward(ctx, g):
        nx, ny, nz = ctx.dims
        gs = F.scalar_storage(g)
        gf = F.new_pop(nx, ny, nz, gs.dtype, gs.device)
        _launch("lbm_density_vjp", F.dtype_tag(gs.dtype), L.ptr(gs), L.ptr(gf), nx, ny, nz,
                L.stream_handle(gs.device))
        return F.pop_view(gf)


class _Velocity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, rho):
        out = _compute_velocity_fwd(f, rho)
        ctx.save_for_backward(f, rho if isinstance(rho, torch.Tensor) else None)
        ctx.rho_like = rho
        return out

    @staticmethod
    def backward(ctx, g):
        f, rho_t = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        rs = F.as_field(rho_t if rho_t is not None else ctx.rho_like, fs, (nx, ny, nz))
        gu = F.vec_storage(g.to(fs.dtype))
        gf = torch.empty_like(fs)
        grho = torch.empty_like(rs) if _needs(ctx.rho_like) else None
        _launch("lbm_velocity_vjp", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(rs), L.ptr(gu), L.ptr(gf), L.ptr(grho),
                nx, ny, nz, L.stream_handle(fs.device))
        return F.pop_view(gf), (_field_grad(grho, rho_t) if grho is not None else None)


class _Equilibrium(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, u):
        out = _equilibrium_fwd(rho, u)
        ctx.save_for_backward(rho if isinstance(rho, torch.Tensor) else None, u)
        ctx.rho_like = rho
        return out

    @staticmethod
    def backward(ctx, g):
        rho_t, u = ctx.saved_tensors
        nx, ny, nz = int(u.shape[1]), int(u.shape[2]), int(u.shape[3])
        dt = g.dtype
        us = F.vec_storage(u.to(dt))
        rs = F.as_field(rho_t if rho_t is not None else ctx.rho_like, us, (nx, ny, nz))
        gs = _pop_grad(g, dt)
        grho = torch.empty_like(rs)
        gu = torch.empty_like(us)
        _launch("lbm_equilibrium_vjp", F.dtype_tag(dt), L.ptr(rs), L.ptr(us), L.ptr(gs), L.ptr(grho), L.ptr(gu),
                nx, ny, nz, L.stream_handle(us.device))
        g_rho = _field_grad(grho, rho_t) if _needs(ctx.rho_like) else None
        g_u = F.vec_view(gu).to(u.dtype) if u.requires_grad else None
        return g_rho, g_u


class _Collide(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, tau, kind):
        out = _collide("lbm_mrt" if kind == L.LBM_COLL_MRT else "lbm_bgk", f, tau)
        ctx.save_for_backward(f, tau if isinstance(tau, torch.Tensor) else None)
        ctx.tau_like, ctx.kind = tau, kind
        return out

    @staticmethod
    def backward(ctx, g):
        f, tau_t = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        tau = tau_t if tau_t is not None else ctx.tau_like
        tau_field, tau_s = None, 0.0
        if isinstance(tau, torch.Tensor) and tau.dim() > 0:
            tau_field = F.as_field(tau, fs, (nx, ny, nz))
        else:
            tau_s = F.scalar_float(tau)
        gs = _pop_grad(g, fs.dtype)
        gf = torch.empty_like(fs)
        gtau = F.new_scalar(nx, ny, nz, fs.dtype, fs.device) if _needs(ctx.tau_like) else None
        _launch("lbm_collide_vjp", F.dtype_tag(fs.dtype), int(ctx.kind), L.ptr(fs), tau_s, L.ptr(tau_field),
                L.ptr(gs), L.ptr(gf), L.ptr(gtau), nx, ny, nz, L.stream_handle(fs.device))
        return F.pop_view(gf), (_field_grad(gtau, tau_t) if gtau is not None else None), None


class _Smagorinsky(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, tau, cs):
        out = _smagorinsky_tau_fwd(f, tau, cs)
        ctx.save_for_backward(f, tau if isinstance(tau, torch.Tensor) else None)
        ctx.tau_like, ctx.cs = tau, float(cs)
        return out

    @staticmethod
    def backward(ctx, g):
        f, tau_t = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        gt = F.scalar_storage(g.to(fs.dtype))
        gf = torch.empty_like(fs)
        gtau = torch.empty_like(gt) if _needs(ctx.tau_like) else None
        _launch("lbm_smagorinsky_tau_vjp", F.dtype_tag(fs.dtype), L.ptr(fs),
                F.scalar_float(tau_t if tau_t is not None else ctx.tau_like), ctx.cs, L.ptr(gt), L.ptr(gf),
                L.ptr(gtau), nx, ny, nz, L.stream_handle(fs.device))
        return F.pop_view(gf), (_field_grad(gtau, tau_t) if gtau is not None else None), None


class _Stream(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, q, periodic_yz):
        out = _stream_fwd(f, q, periodic_yz)
        ctx.save_for_backward(f, q)
        ctx.periodic = int(bool(periodic_yz))
        return out

    @staticmethod
    def backward(ctx, g):
        f, q = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        qs = F.pop_storage(q.to(fs.dtype))
        gs = _pop_grad(g, fs.dtype)
        gf = torch.empty_like(fs)
        gq = torch.empty_like(fs) if q.requires_grad else None
        _launch("lbm_stream_vjp", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(qs), ctx.periodic, L.ptr(gs), L.ptr(gf),
                L.ptr(gq), nx, ny, nz, L.stream_handle(fs.device))
        return F.pop_view(gf), (F.pop_view(gq).to(q.dtype) if gq is not None else None), None


class _ZouHeInlet(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, velocity, rho_target):
        out = _zou_he_inlet_fwd(f, velocity, rho_target)
        ctx.save_for_backward(f)
        ctx.vel = _vec3(velocity)
        ctx.vel_like, ctx.rt_like = velocity, rho_target
        return out

    @staticmethod
    def backward(ctx, g):
        (f,) = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        gs = _pop_grad(g, fs.dtype).clone()
        rows = ny * nz
        want_u, want_rt = _needs(ctx.vel_like), _needs(ctx.rt_like)
        gu = torch.empty((rows, 3), dtype=torch.float64, device=fs.device) if want_u else None
        grt = torch.empty(rows, dtype=torch.float64, device=fs.device) if want_rt else None
        has_rt = ctx.rt_like is not None
        rt = F.scalar_float(ctx.rt_like) if has_rt else 0.0
        _launch("lbm_zou_he_inlet_vjp", F.dtype_tag(fs.dtype), L.ptr(fs), *ctx.vel, int(has_rt), rt, L.ptr(gs),
                L.ptr(gu), L.ptr(grt), nx, ny, nz, L.stream_handle(fs.device))
        # returned on the input's own device (a host velocity tensor is allowed)
        g_u = gu.sum(0).reshape(ctx.vel_like.shape).to(ctx.vel_like.device, ctx.vel_like.dtype) if want_u else None
        g_rt = grt.sum().reshape(ctx.rt_like.shape).to(ctx.rt_like.device, ctx.rt_like.dtype) if want_rt else None
        return F.pop_view(gs), g_u, g_rt


class _ZouHeOutlet(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, rho_out):
        out = _zou_he_outlet_fwd(f, rho_out)
        ctx.dims, ctx.dtype = F.dims_of_pop(f), f.dtype
        ctx.rho_like = rho_out
        return out

    @staticmethod
    def backward(ctx, g):
        nx, ny, nz = ctx.dims
        gs = _pop_grad(g, ctx.dtype).clone()
        want = _needs(ctx.rho_like)
        grho = torch.empty(ny * nz, dtype=torch.float64, device=gs.device) if want else None
        _launch("lbm_zou_he_outlet_vjp", F.dtype_tag(gs.dtype), F.scalar_float(ctx.rho_like), L.ptr(gs), L.ptr(grho),
                nx, ny, nz, L.stream_handle(gs.device))
        g_rho = grho.sum().reshape(ctx.rho_like.shape).to(ctx.rho_like.device, ctx.rho_like.dtype) if want else None
        return F.pop_view(gs), g_rho


class _Momentum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f_post, f_streamed, q, decomposed):
        out, dt = _mex(f_post, f_streamed, q, decomposed)
        ctx.save_for_backward(f_post, q)
        ctx.decomposed, ctx.dt = decomposed, dt
        ctx.post_dtype, ctx.str_dtype = f_post.dtype, f_streamed.dtype
        if decomposed:
            return out[:3].to(dt), out[3:].to(dt)
        return out[:3].to(dt)

    @staticmethod
    def backward(ctx, *grads):
        f_post, q = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f_post)
        ps = F.pop_storage(f_post)
        qs = F.pop_storage(q.to(ps.dtype))
        dev = ps.device
        g_tot = grads[0]
        g_tot = (torch.zeros(3, dtype=torch.float64, device=dev) if g_tot is None
                 else g_tot.detach().to(torch.float64).contiguous())
        g_pr = None
        if ctx.decomposed and grads[1] is not None:
            g_pr = grads[1].detach().to(torch.float64).contiguous()
        gp = torch.empty_like(ps)
        gsr = torch.empty_like(ps)
        _launch("lbm_momentum_exchange_vjp", F.dtype_tag(ps.dtype), L.ptr(ps), L.ptr(qs), L.ptr(g_tot), L.ptr(g_pr),
                L.ptr(gp), L.ptr(gsr), nx, ny, nz, L.stream_handle(dev))
        return F.pop_view(gp).to(ctx.post_dtype), F.pop_view(gsr).to(ctx.str_dtype), None, None


class _Drag(torch.autograd.Function):
    @staticmethod
    def forward(ctx, force, inlet_velocity, area, rho):
        out = _drag_coefficient_fwd(force, inlet_velocity, area, rho)
        ctx.save_for_backward(force)
        ctx.u, ctx.area, ctx.rho = F.scalar_float(inlet_velocity), area, float(rho)
        return out

    @staticmethod
    def backward(ctx, g):
        (force,) = ctx.saved_tensors
        fd = force.detach().to(torch.float64).reshape(-1, 3).contiguous()
        if isinstance(ctx.area, torch.Tensor):
            a = ctx.area.detach().to(device=fd.device, dtype=torch.float64).reshape(1).contiguous()
        else:
            a = torch.full((1,), float(ctx.area), dtype=torch.float64, device=fd.device)
        gc = g.detach().to(torch.float64).reshape(-1).contiguous()
        gfo = torch.empty_like(fd)
        _launch("lbm_drag_coefficient_vjp", L.ptr(fd), L.ptr(gc), int(fd.shape[0]), ctx.u, L.ptr(a), ctx.rho,
                L.ptr(gfo), L.stream_handle(fd.device))
        return gfo.reshape(force.shape).to(force.dtype), None, None, None


class _SoftBounce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, porosity):
        out = _soft_bounce_back_fwd(f, porosity)
        ctx.save_for_backward(f, porosity if isinstance(porosity, torch.Tensor) else None)
        ctx.por_like = porosity
        return out

    @staticmethod
    def backward(ctx, g):
        f, por_t = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        ps = F.as_field(por_t if por_t is not None else ctx.por_like, fs, (nx, ny, nz))
        gs = _pop_grad(g, fs.dtype)
        gf = torch.empty_like(fs)
        gphi = torch.empty_like(ps) if _needs(ctx.por_like) else None
        _launch("lbm_soft_bounce_back_vjp", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(ps), L.ptr(gs), L.ptr(gf),
                L.ptr(gphi), nx, ny, nz, L.stream_handle(fs.device))
        return F.pop_view(gf), (_field_grad(gphi, por_t) if gphi is not None else None)


class _HardBounce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, solid):
        out = _hard_bounce_back_fwd(f, solid)
        ctx.save_for_backward(solid)
        ctx.dims, ctx.dtype = F.dims_of_pop(f), f.dtype
        return out

    @staticmethod
    def backward(ctx, g):
        (solid,) = ctx.saved_tensors
        nx, ny, nz = ctx.dims
        ss = F.scalar_storage(solid.to(torch.int32))
        gs = _pop_grad(g, ctx.dtype)
        gf = torch.empty_like(gs)
        _launch("lbm_hard_bounce_vjp", F.dtype_tag(gs.dtype), L.ptr(ss), L.ptr(gs), L.ptr(gf), nx, ny, nz,
                L.stream_handle(gs.device))
        return F.pop_view(gf), None


class _Guard(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f):
        ctx.save_for_backward(f)
        return _stability_guard_fwd(f)

    @staticmethod
    def backward(ctx, g):
        (f,) = ctx.saved_tensors
        nx, ny, nz = F.dims_of_pop(f)
        fs = F.pop_storage(f)
        gs = _pop_grad(g, fs.dtype)
        gf = torch.empty_like(fs)
        _launch("lbm_stability_guard_vjp", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(gs), L.ptr(gf), nx, ny, nz,
                L.stream_handle(fs.device))
        return F.pop_view(gf).to(f.dtype)


def compute_density(f: torch.Tensor) -> torch.Tensor:
    """rho = sum_i f_i.  jax_lbm/core/equilibrium.py:22-48."""
    return _Density.apply(f)


def compute_velocity(f: torch.Tensor, rho: torch.Tensor) -> torch.Tensor:
    """u = sum_i f_i e_i / rho.  jax_lbm/core/equilibrium.py:51-87."""
    return _Velocity.apply(f, rho)


def equilibrium(rho, u: torch.Tensor) -> torch.Tensor:
    """Second-order D3Q19 equilibrium.  jax_lbm/core/equilibrium.py:90-151."""
    return _Equilibrium.apply(rho, u)


def bgk(f: torch.Tensor, tau) -> torch.Tensor:
    """BGK collision; tau scalar or (nx, ny, nz).  jax_lbm/core/collision.py:46-100."""
    return _Collide.apply(f, tau, L.LBM_COLL_BGK)


def mrt(f: torch.Tensor, tau) -> torch.Tensor:
    """MRT collision (d'Humieres).  jax_lbm/core/collision.py:103-184."""
    return _Collide.apply(f, tau, L.LBM_COLL_MRT)


def smagorinsky_tau(f: torch.Tensor, tau, cs: float = 0.1) -> torch.Tensor:
    """Effective relaxation time, shape (nx, ny, nz).  jax_lbm/core/collision.py:187-270.
    Reverse mode takes the square root's derivative at |Pi_neq| = 0 as 0 (JAX gives NaN)."""
    return _Smagorinsky.apply(f, tau, cs)


def stream(f: torch.Tensor, q: torch.Tensor, periodic_yz: bool = True) -> torch.Tensor:
    """Pull streaming with Bouzidi links.  jax_lbm/core/streaming.py:40-185."""
    return _Stream.apply(f, q, periodic_yz)


def zou_he_inlet(f: torch.Tensor, velocity, rho_target=None) -> torch.Tensor:
    """Zou-He velocity inlet at x = 0.  jax_lbm/core/boundary.py:35-124."""
    return _ZouHeInlet.apply(f, velocity, rho_target)


def zou_he_outlet(f: torch.Tensor, rho_out: float = 1.0) -> torch.Tensor:
    """Zou-He pressure outlet at x = nx-1.  jax_lbm/core/boundary.py:127-200."""
    return _ZouHeOutlet.apply(f, rho_out)


def momentum_exchange(f_post: torch.Tensor, f_streamed: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """Force on the walls, (Fx, Fy, Fz), accumulated in float64.  jax_lbm/core/forces.py:37-120.
    Differentiable in f_post and f_streamed; q is a mask (no cotangent), as in the reference."""
    return _Momentum.apply(f_post, f_streamed, q, False)


def momentum_exchange_decomposed(f_post, f_streamed, q):
    """(total, pressure) forces.  jax_lbm/core/forces.py:123-196."""
    return _Momentum.apply(f_post, f_streamed, q, True)


def drag_coefficient(force: torch.Tensor, inlet_velocity: float, area, rho: float = 1.0) -> torch.Tensor:
    """Cd = |F_x| / (0.5 rho U^2 A + 1e-10).  jax_lbm/core/forces.py:237-276.

    ``force`` may be (3,) or a (T, 3) history; the result is () or (T,).
    Differentiable in ``force``.
    """
    return _Drag.apply(force, inlet_velocity, area, rho)


def soft_bounce_back(f_collided: torch.Tensor, porosity: torch.Tensor) -> torch.Tensor:
    """(1 - phi) f + phi f[opp].  jax_lbm/geometry/smooth.py:118-155."""
    return _SoftBounce.apply(f_collided, porosity)


def hard_bounce_back(f_collided: torch.Tensor, solid: torch.Tensor) -> torch.Tensor:
    """f[opp] on cells with solid >= 1 (jax_lbm/state.py:146-149)."""
    return _HardBounce.apply(f_collided, solid)


def stability_guard(f: torch.Tensor) -> torch.Tensor:
    """Cells with rho < 0.3, rho > 2, |u| > 0.4 or rho = NaN become f_eq(1, 0);
    the guard that ends step() at jax_lbm/state.py:158-169.  Reset cells pass
    no cotangent (the reset value is a constant)."""
    return _Guard.apply(f)
