"""LBMState / LBMParams / step / simulate (jax_lbm/state.py:33-230) on the fused kernels.

``step`` runs the reference's canonical order — [Smagorinsky] -> BGK|MRT ->
hard bounce (solid >= 1) -> stream with Bouzidi -> Zou-He inlet -> Zou-He
outlet -> post-step stability guard (state.py:129-171) — as two launches of
the fused kernel (collide, then stream).  ``simulate`` runs n steps with one
launch per step.

Reverse mode (the reference differentiates step/simulate with jax.grad
through every LBMParams leaf):
* when the populations, q or a tensor inlet velocity require grad, with
  periodic y/z and a scalar tau that does not, both run through the
  differentiable fused rollout (hand-written adjoint of the whole step);
* when tau requires grad or is a per-cell field, or y/z are clamped
  (periodic_yz=False, streaming.py:134-142), they run the same step as a
  composition of the per-operator kernels, each with its hand-written VJP
  kernel (and a per-cell tau runs that way forward too, since the fused
  step takes one tau).
``checkpoint`` picks the memory policy, as jax.checkpoint does in the
reference: True recomputes segments of about sqrt(n) steps, False keeps
every state.
"""

from __future__ import annotations

import math
from typing import NamedTuple

import torch
from torch.utils.checkpoint import checkpoint as _recompute

from . import _field as F
from .core import ops as _ops
from .engine import Lattice


class LBMState(NamedTuple):
    """f (19, nx, ny, nz), solid (nx, ny, nz) int32, q (19, nx, ny, nz).  state.py:33-54."""

    f: torch.Tensor
    solid: torch.Tensor
    q: torch.Tensor


class LBMParams(NamedTuple):
    """tau, inlet_velocity (3,), collision_mode 0=BGK 1=MRT, use_smagorinsky,
    smagorinsky_cs, periodic_yz.  state.py:57-84."""

    tau: object
    inlet_velocity: object
    collision_mode: int = 0
    use_smagorinsky: int = 0
    smagorinsky_cs: float = 0.1
    periodic_yz: bool = True


# The lattice of the last forward step/simulate call, so that a loop of
# step() calls on one geometry compiles its link list once.  The entry holds
# the solid and q tensors themselves and their versions: it serves only those
# objects (never a later tensor at a reused address) and only unchanged (an
# in-place write bumps the version and rebuilds it).
_last_lattice = None


def _version(t):
    return None if t is None else t._version


def _lattice(state: LBMState, params: LBMParams) -> Lattice:
    global _last_lattice
    nx, ny, nz = F.dims_of_pop(state.f)
    vel = params.inlet_velocity
    if isinstance(vel, torch.Tensor):
        vel = vel.detach().reshape(-1).to("cpu", torch.float64).tolist()
    vel = tuple(float(v) for v in vel)
    storage = "f64" if state.f.dtype == torch.float64 else "f32"
    key = (nx, ny, nz, str(state.f.device), storage, F.scalar_float(params.tau), vel, int(params.collision_mode),
           int(params.use_smagorinsky), float(params.smagorinsky_cs), bool(params.periodic_yz))
    if _last_lattice is not None:
        k, solid, q, versions, lat = _last_lattice
        if (k == key and solid is state.solid and q is state.q
                and versions == (_version(state.solid), _version(state.q))):
            return lat
    lat = Lattice(nx, ny, nz, tau=F.scalar_float(params.tau), inlet_velocity=vel,
                  collision="mrt" if int(params.collision_mode) == 1 else "bgk",
                  smagorinsky_cs=float(params.smagorinsky_cs) if int(params.use_smagorinsky) else None,
                  solid=state.solid, q=state.q, guard="post", periodic_yz=bool(params.periodic_yz),
                  storage=storage, device=state.f.device)
    _last_lattice = (key, state.solid, state.q, (_version(state.solid), _version(state.q)), lat)
    return lat


def _tau_is_field(params: LBMParams) -> bool:
    return isinstance(params.tau, torch.Tensor) and params.tau.dim() > 0 and params.tau.numel() > 1


def _tau_needs_grad(params: LBMParams) -> bool:
    return isinstance(params.tau, torch.Tensor) and params.tau.requires_grad


def _wants_grad(state: LBMState, params: LBMParams) -> bool:
    if not torch.is_grad_enabled():
        return False
    vel = params.inlet_velocity
    return bool(state.f.requires_grad or (state.q is not None and state.q.requires_grad)
                or (isinstance(vel, torch.Tensor) and vel.requires_grad) or _tau_needs_grad(params))


def _per_op_path(state: LBMState, params: LBMParams, grad: bool) -> bool:
    """Per-operator composition: a per-cell tau (any mode), or in reverse mode
    a tau that requires grad or clamped y/z (the fused adjoint covers the rest)."""
    if _tau_is_field(params):
        return True
    return grad and (_tau_needs_grad(params) or not bool(params.periodic_yz))


def _step_ops(f: torch.Tensor, state: LBMState, params: LBMParams) -> torch.Tensor:
    """One step as differentiable per-operator kernels, in the order of
    jax_lbm/state.py:129-169: [Smagorinsky tau] -> BGK|MRT -> hard bounce
    (solid >= 1) -> stream with Bouzidi -> Zou-He inlet -> Zou-He outlet ->
    stability guard."""
    tau = params.tau
    if int(params.use_smagorinsky):
        tau = _ops.smagorinsky_tau(f, tau, float(params.smagorinsky_cs))
    f = _ops.mrt(f, tau) if int(params.collision_mode) == 1 else _ops.bgk(f, tau)
    f = _ops.hard_bounce_back(f, state.solid)
    f = _ops.stream(f, state.q, bool(params.periodic_yz))
    f = _ops.zou_he_inlet(f, params.inlet_velocity)
    f = _ops.zou_he_outlet(f, 1.0)
    return _ops.stability_guard(f)


def _ops_rollout(state: LBMState, params: LBMParams, n: int, checkpoint: bool) -> torch.Tensor:
    """n steps of ``_step_ops``; with checkpoint each segment of about sqrt(n)
    steps keeps only its input and is recomputed in the backward."""

    def segment(f, k):
        for _ in range(k):
            f = _step_ops(f, state, params)
        return f

    seg = max(1, int(math.ceil(math.sqrt(n)))) if checkpoint and n > 1 else n
    f, t = state.f, 0
    while t < n:
        k = min(seg, n - t)
        if seg < n and torch.is_grad_enabled():
            f = _recompute(segment, f, k, use_reentrant=False)
        else:
            f = segment(f, k)
        t += k
    return f


def _rollout(state: LBMState, params: LBMParams, n: int, checkpoint: bool) -> torch.Tensor:
    from .adjoint import rollout
    vel = params.inlet_velocity
    if not isinstance(vel, torch.Tensor):
        vel = [float(v) for v in vel]
    policy = max(1, int(math.ceil(math.sqrt(n)))) if checkpoint and n > 1 else "all"
    storage = "f64" if state.f.dtype == torch.float64 else "f32"
    S, _, _ = rollout(state.f, n, tau=F.scalar_float(params.tau), q=state.q, solid=state.solid,
                      inlet_velocity=vel, collision="mrt" if int(params.collision_mode) == 1 else "bgk",
                      smagorinsky_cs=float(params.smagorinsky_cs) if int(params.use_smagorinsky) else None,
                      guard="post", storage=storage, checkpoint=policy)
    return S


def step(state: LBMState, params: LBMParams) -> LBMState:
    """One timestep.  jax_lbm/state.py:87-171."""
    F.require_cuda(state.f, "state.f")
    grad = _wants_grad(state, params)
    if _per_op_path(state, params, grad):
        return LBMState(_ops_rollout(state, params, 1, False), state.solid, state.q)
    if grad:
        return LBMState(_rollout(state, params, 1, False), state.solid, state.q)
    lat = _lattice(state, params)
    s_final, _, _ = lat.run(state.f, 1, forces=False)
    return LBMState(s_final, state.solid, state.q)


def simulate(state: LBMState, params: LBMParams, n_steps: int, checkpoint: bool = True) -> LBMState:
    """n_steps timesteps.  jax_lbm/state.py:174-230."""
    F.require_cuda(state.f, "state.f")
    if int(n_steps) == 0:
        return state
    grad = _wants_grad(state, params)
    if _per_op_path(state, params, grad):
        return LBMState(_ops_rollout(state, params, int(n_steps), checkpoint), state.solid, state.q)
    if grad:
        return LBMState(_rollout(state, params, int(n_steps), checkpoint), state.solid, state.q)
    lat = _lattice(state, params)
    s_final, _, _ = lat.run(state.f, int(n_steps), forces=False)
    return LBMState(s_final, state.solid, state.q)
