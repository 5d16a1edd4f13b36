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


def compute_density(f: torch.Tensor) -> torch.Tensor:
    """rho = sum_i f_i.  jax_lbm/core/equilibrium.py:22-48."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    out = F.new_scalar(nx, ny, nz, fs.dtype, fs.device)
    _launch("lbm_density", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(out), nx, ny, nz, L.stream_handle(fs.device))
    return F.scalar_view(out)


def compute_velocity(f: torch.Tensor, rho: torch.Tensor) -> torch.Tensor:
    """u = sum_i f_i e_i / rho.  jax_lbm/core/equilibrium.py:51-87."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    rs = F.as_field(rho, fs, (nx, ny, nz))
    out = torch.empty((3, nz, ny, nx), dtype=fs.dtype, device=fs.device)
    _launch("lbm_velocity", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(rs), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.vec_view(out)


def equilibrium(rho, u: torch.Tensor) -> torch.Tensor:
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


def bgk(f: torch.Tensor, tau) -> torch.Tensor:
    """BGK collision; tau scalar or (nx, ny, nz).  jax_lbm/core/collision.py:46-100."""
    return _collide("lbm_bgk", f, tau)


def mrt(f: torch.Tensor, tau) -> torch.Tensor:
    """MRT collision (d'Humieres).  jax_lbm/core/collision.py:103-184."""
    return _collide("lbm_mrt", f, tau)


def smagorinsky_tau(f: torch.Tensor, tau, cs: float = 0.1) -> torch.Tensor:
    """Effective relaxation time, shape (nx, ny, nz).  jax_lbm/core/collision.py:187-270."""
    F.require_cuda(f, "f")
    nx, ny, nz = F.dims_of_pop(f)
    fs = F.pop_storage(f)
    out = F.new_scalar(nx, ny, nz, fs.dtype, fs.device)
    _launch("lbm_smagorinsky_tau", F.dtype_tag(fs.dtype), L.ptr(fs), F.scalar_float(tau), float(cs), L.ptr(out),
            nx, ny, nz, L.stream_handle(fs.device))
    return F.scalar_view(out)


def stream(f: torch.Tensor, q: torch.Tensor, periodic_yz: bool = True) -> torch.Tensor:
    """Pull streaming with Bouzidi links.  jax_lbm/core/streaming.py:40-185."""
    F.require_cuda(f, "f")
    F.require_cuda(q, "q")
    nx, ny, nz = F.dims_of_pop(f)
    if tuple(q.shape) != tuple(f.shape):
        raise ValueError(f"q must have the shape of f {tuple(f.shape)}, got {tuple(q.shape)}")
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


def zou_he_inlet(f: torch.Tensor, velocity, rho_target=None) -> torch.Tensor:
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


def zou_he_outlet(f: torch.Tensor, rho_out: float = 1.0) -> torch.Tensor:
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
    ps = F.pop_storage(f_post)
    ss = F.pop_storage(f_streamed.to(ps.dtype))
    qs = F.pop_storage(q.to(ps.dtype))
    ws = torch.empty(int(L.load().lbm_reduce_workspace_bytes(nx, ny, nz)) // 8, dtype=torch.float64,
                     device=ps.device)
    out = torch.empty(6, dtype=torch.float64, device=ps.device)
    _launch("lbm_momentum_exchange", F.dtype_tag(ps.dtype), L.ptr(ps), L.ptr(ss), L.ptr(qs), int(want_pressure),
            L.ptr(out), L.ptr(ws), nx, ny, nz, L.stream_handle(ps.device))
    return out, ps.dtype


def momentum_exchange(f_post: torch.Tensor, f_streamed: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """Force on the walls, (Fx, Fy, Fz), accumulated in float64.  jax_lbm/core/forces.py:37-120."""
    out, dt = _mex(f_post, f_streamed, q, False)
    return out[:3].to(dt)


def momentum_exchange_decomposed(f_post, f_streamed, q):
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


def drag_coefficient(force: torch.Tensor, inlet_velocity: float, area, rho: float = 1.0) -> torch.Tensor:
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


def soft_bounce_back(f_collided: torch.Tensor, porosity: torch.Tensor) -> torch.Tensor:
    """(1 - phi) f + phi f[opp].  jax_lbm/geometry/smooth.py:118-155."""
    F.require_cuda(f_collided, "f_collided")
    nx, ny, nz = F.dims_of_pop(f_collided)
    fs = F.pop_storage(f_collided)
    ps = F.as_field(porosity, fs, (nx, ny, nz))
    out = torch.empty_like(fs)
    _launch("lbm_soft_bounce_back", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(ps), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.pop_view(out)


def hard_bounce_back(f_collided: torch.Tensor, solid: torch.Tensor) -> torch.Tensor:
    """f[opp] on cells with solid >= 1 (jax_lbm/state.py:146-149)."""
    F.require_cuda(f_collided, "f_collided")
    nx, ny, nz = F.dims_of_pop(f_collided)
    fs = F.pop_storage(f_collided)
    ss = F.scalar_storage(solid.to(torch.int32))
    out = torch.empty_like(fs)
    _launch("lbm_hard_bounce", F.dtype_tag(fs.dtype), L.ptr(fs), L.ptr(ss), L.ptr(out), nx, ny, nz,
            L.stream_handle(fs.device))
    return F.pop_view(out)
