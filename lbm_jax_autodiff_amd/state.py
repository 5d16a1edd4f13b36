"""LBMState / LBMParams / step / simulate (jax_lbm/state.py:33-230) on the fused kernels.

``step`` runs the reference's canonical order — [Smagorinsky] -> BGK|MRT ->
hard bounce (solid >= 1) -> stream with Bouzidi -> Zou-He inlet -> Zou-He
outlet -> post-step stability guard (state.py:129-171) — as two launches of
the fused kernel (collide, then stream).  ``simulate`` runs n steps with one
launch per step.

When the populations, q or a tensor inlet velocity require grad, both run
through the differentiable fused rollout (hand-written adjoint), with
periodic y/z as the adjoint requires.  ``checkpoint`` then picks its memory
policy, as jax.checkpoint does in the reference: True recomputes segments of
about sqrt(n) steps, False keeps every state.
"""

from __future__ import annotations

import math
from typing import NamedTuple

import torch

from . import _field as F
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


def _lattice(state: LBMState, params: LBMParams) -> Lattice:
    nx, ny, nz = F.dims_of_pop(state.f)
    vel = params.inlet_velocity
    if isinstance(vel, torch.Tensor):
        vel = vel.detach().reshape(-1).to("cpu", torch.float64).tolist()
    storage = "f64" if state.f.dtype == torch.float64 else "f32"
    return Lattice(nx, ny, nz, tau=F.scalar_float(params.tau), inlet_velocity=vel,
                   collision="mrt" if int(params.collision_mode) == 1 else "bgk",
                   smagorinsky_cs=float(params.smagorinsky_cs) if int(params.use_smagorinsky) else None,
                   solid=state.solid, q=state.q, guard="post", periodic_yz=bool(params.periodic_yz),
                   storage=storage, device=state.f.device)


def _wants_grad(state: LBMState, params: LBMParams) -> bool:
    if not torch.is_grad_enabled():
        return False
    vel = params.inlet_velocity
    return bool(state.f.requires_grad or (state.q is not None and state.q.requires_grad)
                or (isinstance(vel, torch.Tensor) and vel.requires_grad))


def _rollout(state: LBMState, params: LBMParams, n: int, checkpoint: bool) -> torch.Tensor:
    from .adjoint import rollout
    if not bool(params.periodic_yz):
        raise NotImplementedError("reverse mode of step/simulate needs periodic_yz=True")
    if isinstance(params.tau, torch.Tensor) and params.tau.requires_grad:
        raise NotImplementedError("reverse mode of step/simulate does not differentiate tau")
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
    if _wants_grad(state, params):
        return LBMState(_rollout(state, params, 1, False), state.solid, state.q)
    lat = _lattice(state, params)
    s_final, _, _ = lat.run(state.f, 1, forces=False)
    return LBMState(s_final, state.solid, state.q)


def simulate(state: LBMState, params: LBMParams, n_steps: int, checkpoint: bool = True) -> LBMState:
    """n_steps timesteps.  jax_lbm/state.py:174-230."""
    F.require_cuda(state.f, "state.f")
    if int(n_steps) == 0:
        return state
    if _wants_grad(state, params):
        return LBMState(_rollout(state, params, int(n_steps), checkpoint), state.solid, state.q)
    lat = _lattice(state, params)
    s_final, _, _ = lat.run(state.f, int(n_steps), forces=False)
    return LBMState(s_final, state.solid, state.q)
