"""jax_lbm.geometry equivalents backed by liblbm_hip.so.

Sphere geometry on the reference's world box [-4,4] x [-2,2] x [-2,2] (node
(i,j,k) at ((i+1/2)/(nx/8) - 4, (j+1/2)/(ny/4) - 2, (k+1/2)/(nz/4) - 2),
jax_lbm/geometry/primitives.py:26-53).  ``center``, ``radius`` and
``sharpness`` are in world units.  ``smooth_sphere_geometry`` is
differentiable in ``radius`` (a 0-d tensor with requires_grad): dq/dr and
d porosity/dr come from the same kernel and are chained with float64 dots.
"""

from __future__ import annotations

import torch

from .. import _field as F
from .. import _lib as L
from ..core.ops import soft_bounce_back  # noqa: F401  (re-export, smooth.py:118-155)

__all__ = ["create_sphere", "compute_q_from_solid", "signed_distance_sphere", "porosity_from_sdf",
           "soft_bounce_back", "smooth_sphere_geometry", "sphere_geometry_raw"]


def _device(device):
    return torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())


def _tag(dtype):
    return F.dtype_tag(dtype)


def sphere_geometry_raw(nx, ny, nz, center, radius, *, mode, sharpness=20.0, coord_dtype=torch.float32,
                        out_dtype=torch.float32, disc_floor=1e-10, denom_eps=1e-20, with_grads=False,
                        device=None):
    """Run the sphere geometry kernel; returns lattice-storage tensors
    (solid, porosity, q, dqdr, dphidr) (the last two None unless with_grads)."""
    dev = _device(device)
    N = nx * ny * nz
    solid = torch.empty((nz, ny, nx), dtype=torch.int32, device=dev)
    porosity = torch.empty((nz, ny, nx), dtype=out_dtype, device=dev)
    q = torch.empty((19, nz, ny, nx), dtype=out_dtype, device=dev)
    dqdr = torch.empty_like(q) if with_grads else None
    dphidr = torch.empty_like(porosity) if with_grads else None
    cx, cy, cz = (float(c) for c in center)
    L.call("lbm_sphere_geometry", int(mode), _tag(coord_dtype), _tag(out_dtype), cx, cy, cz, F.scalar_float(radius),
           float(sharpness), float(disc_floor), float(denom_eps), L.ptr(solid), L.ptr(porosity), L.ptr(q),
           L.ptr(dqdr), L.ptr(dphidr), nx, ny, nz, L.stream_handle(dev))
    del N
    return solid, porosity, q, dqdr, dphidr


def dot64(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Float64 fixed-order sum of a*b on the device (0-d float64 tensor)."""
    a = a.contiguous()
    b = b.contiguous()
    if a.numel() != b.numel():
        raise ValueError("dot64: size mismatch")
    ws = torch.empty(int(L.load().lbm_reduce_workspace_bytes(1, 1, 1)) // 8, dtype=torch.float64, device=a.device)
    out = torch.empty(1, dtype=torch.float64, device=a.device)
    L.call("lbm_dot", _tag(a.dtype), L.ptr(a), _tag(b.dtype), L.ptr(b), a.numel(), L.ptr(out), L.ptr(ws),
           L.stream_handle(a.device))
    return out.reshape(())


def create_sphere(nx, ny, nz, center, radius, device=None):
    """Hard sphere (dist^2 <= r^2) and ray-sphere Bouzidi q.
    jax_lbm/geometry/primitives.py:56-174.  Returns (solid int32, q float32);
    q is differentiable in ``radius`` and ``center`` when they require grad."""
    dev = device
    for v in (radius, center):
        if isinstance(v, torch.Tensor) and v.is_cuda:
            dev = v.device
    dev = _device(dev)
    if _needs_grad(radius) or _needs_grad(center):
        _, solid, q = _SphereGeometry.apply(radius, center, nx, ny, nz, 0, 0.0, torch.float32, torch.float32,
                                            0.0, 1e-20, dev)
        return solid, q
    solid, _, q, _, _ = sphere_geometry_raw(nx, ny, nz, _center_floats(center), radius, mode=0, device=dev)
    return F.scalar_view(solid), F.pop_view(q)


def compute_q_from_solid(solid: torch.Tensor, nx, ny, nz, default_q: float = 0.5, dtype=torch.float32):
    """q = default_q on fluid -> body (solid == 1) in-bounds links, else -1.
    jax_lbm/geometry/bouzidi.py:32-106."""
    F.require_cuda(solid, "solid")
    ss = F.scalar_storage(solid.to(torch.int32))
    q = torch.empty((19, nz, ny, nx), dtype=dtype, device=ss.device)
    L.call("lbm_q_from_solid", _tag(dtype), L.ptr(ss), float(default_q), L.ptr(q), nx, ny, nz,
           L.stream_handle(ss.device))
    return F.pop_view(q)


def signed_distance_sphere(nx, ny, nz, center, radius, device=None, coord_dtype=torch.float32,
                           out_dtype=torch.float32):
    """r - |x_w - c|, positive inside.  jax_lbm/geometry/smooth.py:34-78."""
    dev = _device(device)
    out = torch.empty((nz, ny, nx), dtype=out_dtype, device=dev)
    cx, cy, cz = (float(c) for c in center)
    L.call("lbm_sphere_sdf", _tag(coord_dtype), _tag(out_dtype), cx, cy, cz, F.scalar_float(radius), L.ptr(out),
           nx, ny, nz, L.stream_handle(dev))
    return F.scalar_view(out)


def porosity_from_sdf(sdf: torch.Tensor, sharpness: float = 20.0) -> torch.Tensor:
    """sigmoid(sharpness * sdf).  jax_lbm/geometry/smooth.py:81-115."""
    F.require_cuda(sdf, "sdf")
    nx, ny, nz = (int(s) for s in sdf.shape)
    ss = F.scalar_storage(sdf)
    out = torch.empty_like(ss)
    L.call("lbm_porosity_from_sdf", _tag(ss.dtype), L.ptr(ss), float(sharpness), L.ptr(out), nx, ny, nz,
           L.stream_handle(ss.device))
    return F.scalar_view(out)


def _center_floats(center):
    if isinstance(center, torch.Tensor):
        vals = center.detach().reshape(-1).to("cpu", torch.float64).tolist()
    else:
        vals = [float(c) for c in center]
    if len(vals) != 3:
        raise ValueError("center must have 3 components")
    return tuple(vals)


class _SphereGeometry(torch.autograd.Function):
    """(porosity, solid, q) of a sphere as differentiable functions of the
    radius and the centre; the backward is one fixed-order reduction kernel
    (lbm_sphere_geometry_vjp), so no per-cell derivative fields are stored."""

    @staticmethod
    def forward(ctx, radius, center, nx, ny, nz, mode, sharpness, coord_dtype, out_dtype, disc_floor, denom_eps,
                device):
        # an output the loss does not use arrives as None, not as zeros: without
        # a porosity cotangent the backward sweeps only the sphere's box
        ctx.set_materialize_grads(False)
        c = _center_floats(center)
        solid, por, q, _, _ = sphere_geometry_raw(nx, ny, nz, c, radius, mode=mode, sharpness=sharpness,
                                                  coord_dtype=coord_dtype, out_dtype=out_dtype,
                                                  disc_floor=disc_floor, denom_eps=denom_eps, device=device)
        ctx.save_for_backward(solid)
        ctx.cfg = (nx, ny, nz, mode, sharpness, coord_dtype, out_dtype, disc_floor, denom_eps, c,
                   F.scalar_float(radius))
        ctx.r_like, ctx.c_like = radius, center
        ctx.mark_non_differentiable(solid)
        return F.scalar_view(por), F.scalar_view(solid), F.pop_view(q)

    @staticmethod
    def backward(ctx, g_por, g_solid, g_q):
        (solid,) = ctx.saved_tensors
        nx, ny, nz, mode, sharpness, coord_dtype, out_dtype, disc_floor, denom_eps, c, r = ctx.cfg
        dev = solid.device
        gp = F.scalar_storage(g_por.to(out_dtype)) if g_por is not None else None
        gq = F.pop_storage(g_q.to(out_dtype)) if g_q is not None else None
        out = torch.empty(4, dtype=torch.float64, device=dev)
        ws = torch.empty(int(L.load().lbm_sphere_geometry_vjp_workspace_bytes()) // 8, dtype=torch.float64,
                         device=dev)
        L.call("lbm_sphere_geometry_vjp", int(mode), _tag(coord_dtype), _tag(out_dtype), c[0], c[1], c[2], r,
               float(sharpness), float(disc_floor), float(denom_eps), L.ptr(solid), L.ptr(gp), L.ptr(gq),
               L.ptr(out), L.ptr(ws), nx, ny, nz, L.stream_handle(dev))
        g_r = g_c = None
        if isinstance(ctx.r_like, torch.Tensor) and ctx.r_like.requires_grad:
            g_r = out[0].reshape(ctx.r_like.shape).to(ctx.r_like.dtype).to(ctx.r_like.device)
        if isinstance(ctx.c_like, torch.Tensor) and ctx.c_like.requires_grad:
            g_c = out[1:].reshape(ctx.c_like.shape).to(ctx.c_like.dtype).to(ctx.c_like.device)
        return (g_r, g_c) + (None,) * 10


def _needs_grad(v) -> bool:
    return isinstance(v, torch.Tensor) and v.requires_grad and torch.is_grad_enabled()


def smooth_sphere_geometry(nx, ny, nz, center, radius, sharpness: float = 20.0, *, coord_dtype=torch.float32,
                           out_dtype=torch.float32, disc_floor=1e-10, denom_eps=1e-20, device=None):
    """(porosity, solid, q) with porosity = sigmoid(sharpness (r - dist)),
    solid = porosity > 1/2 and ray-sphere q(r).  jax_lbm/geometry/smooth.py:158-248.

    ``radius`` and ``center`` (a (3,) tensor) may require grad: porosity and q
    then carry their derivatives (dq = 0 where q is clipped or invalid, as
    JAX's clip/where give; the mask and the link set carry none).
    """
    dev = device
    for v in (radius, center):
        if isinstance(v, torch.Tensor) and v.is_cuda:
            dev = v.device
    dev = _device(dev)
    if _needs_grad(radius) or _needs_grad(center):
        return _SphereGeometry.apply(radius, center, nx, ny, nz, 1, float(sharpness), coord_dtype, out_dtype,
                                     float(disc_floor), float(denom_eps), dev)
    solid, por, q, _, _ = sphere_geometry_raw(nx, ny, nz, _center_floats(center), radius, mode=1,
                                              sharpness=sharpness, coord_dtype=coord_dtype, out_dtype=out_dtype,
                                              disc_floor=disc_floor, denom_eps=denom_eps, device=dev)
    return F.scalar_view(por), F.scalar_view(solid), F.pop_view(q)
