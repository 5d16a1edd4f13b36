"""Reverse mode over a whole rollout with the hand-written HIP adjoint.

``rollout(f0, n_steps, q=..., porosity=..., inlet_velocity=..., ...)`` runs the
fused forward (engine.Lattice) and returns (S_N, P_{N-1}, F[N, 3]) as torch
tensors that are differentiable with respect to ``f0``, the dense Bouzidi
``q``, the soft-bounce ``porosity`` and the inlet velocity.  It replaces
``jax.grad`` / ``jax.value_and_grad`` over the reference's scan loops
(examples/11_optimisation_v2.py:141-176, tests/test_gradients_f64.py:95-217)
and its sqrt(N) ``jax.checkpoint`` chunking (examples/08_gradient_convergence.py:49-66).

Memory policies (``checkpoint``):
  "all"      keep every P_t (store-all);
  K (int)    keep P_t every K steps, recompute each K-step segment once;
  (K1, K2)   two levels: keep every K1, recompute a segment keeping every
             K2, recompute each K2 sub-segment keeping all.
Recomputation is bit-identical to the forward (same kernels, no atomics).
"""

from __future__ import annotations

import ctypes
import math
from typing import Optional, Sequence, Union

import numpy as np
import torch

from . import _field as F
from . import _lib as L
from .engine import Lattice

_E = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [-1, 1, 0],
               [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1], [0, 1, 1], [0, -1, 1],
               [0, 1, -1], [0, -1, -1]], dtype=np.int64)
_OPP = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15], dtype=np.int64)


class AdjointDesc(ctypes.Structure):
    """Mirror of ``lbm_adjoint_desc`` (include/lbm_hip.h)."""

    _fields_ = [
        ("dq_acc", ctypes.c_void_p), ("mu_link", ctypes.c_void_p), ("phi_acc", ctypes.c_void_p),
        ("u_acc", ctypes.c_void_p), ("n_targets", ctypes.c_int32), ("tgt_index", ctypes.c_void_p),
        ("tgt_start", ctypes.c_void_p), ("src", ctypes.c_void_p), ("coef", ctypes.c_void_p),
        ("lam_storage", ctypes.c_int32), ("lam_scale", ctypes.c_double),
        ("mu_link_alt", ctypes.c_void_p), ("fold_links", ctypes.c_int32),
    ]


def scatter_plan(lat: Lattice):
    """Per-entry CSR of the cotangent contributions outside the owner writes.

    For each link (cell x, wall direction j, arriving direction i = opp(j)) of
    the reference's Bouzidi rule (streaming.py:113-181):
      q >= 1/2: S_i(x) = P_j(x)/(2q) + (1 - 1/(2q)) P_i(x)
      q <  1/2: S_i(x) = 2q P_j(x) + (1 - 2q) P_j(x + e_i)   (x + e_i wrapped)
    and the force term F += e_j P_j(x) (forces.py:95-118).  Entries are
    ordered by target and then by source, so the sums are deterministic.
    Built with tensor operations on the links' device (one host sync, for the
    number of wall links).
    """
    n = lat.links.n
    if n == 0:
        return None
    nx, ny, nz = lat.nx, lat.ny, lat.nz
    N = lat.P  # population stride of the cotangent buffers
    RS = lat.RS  # row stride
    dev = lat.links.cell.device
    cell = lat.links.cell.to(torch.int64)
    q = lat.links.q.to(torch.float64)  # (18, n)
    x = cell % nx
    y = (cell // nx) % ny
    z = cell // (nx * ny)
    off = (z * ny + y) * RS + x  # element offset of each link cell in a population
    jm1, b = torch.nonzero(q >= 0.0, as_tuple=True)  # every wall link (j - 1, link index)
    j = jm1 + 1
    i = torch.as_tensor(_OPP, device=dev)[j]
    e = torch.as_tensor(_E, device=dev)
    qq = q[jm1, b]
    cc = off[b]
    src = i * n + b
    hi = qq >= 0.5
    inv2q = 1.0 / (2.0 * qq.clamp(0.5, 1.0))
    twoq = 2.0 * qq.clamp(0.0, 0.5)
    # P_j(x + e_i), the low branch's second neighbour, all axes wrapped
    fx = (x[b] + e[i, 0]) % nx
    fy = (y[b] + e[i, 1]) % ny
    fz = (z[b] + e[i, 2]) % nz
    key = torch.cat([
        j * N + cc,                                                    # own P_j(x): both branches
        torch.where(hi, i * N + cc, j * N + (fz * ny + fy) * RS + fx),  # own P_i(x) | P_j(x + e_i)
        j * N + cc,                                                    # force term on P_j(x)
    ])
    srcs = torch.cat([src, src, -1 - j])
    coef = torch.cat([torch.where(hi, inv2q, twoq), torch.where(hi, 1.0 - inv2q, 1.0 - twoq),
                      torch.zeros_like(qq)])
    # order by target, then by source ((target, source) pairs are unique)
    o1 = torch.argsort(srcs, stable=True)
    o2 = torch.argsort(key[o1], stable=True)
    order = o1[o2]
    key, srcs, coef = key[order], srcs[order], coef[order]
    uniq, counts = torch.unique_consecutive(key, return_counts=True)
    start = torch.cat([torch.zeros(1, dtype=torch.int64, device=dev), torch.cumsum(counts, 0)])
    out = lat.device
    return (uniq.to(torch.int32).to(out), start.to(torch.int32).to(out), srcs.to(torch.int32).to(out),
            coef.to(out))


class _Checkpoints:
    """Yields P_t for t = N-1 down to 0 under a memory policy."""

    def __init__(self, lat: Lattice, policy, n_steps: int):
        self.lat = lat
        self.N = n_steps
        if policy == "all" or policy is None:
            self.levels = [1]
        elif isinstance(policy, int):
            self.levels = [max(1, int(policy)), 1]
        else:
            k1, k2 = (int(v) for v in policy)
            self.levels = [max(1, k1), max(1, min(k2, k1)), 1]
        self.top = {}

    @property
    def stride(self):
        return self.levels[0]

    @property
    def store_all(self):
        return self.levels == [1]

    def record(self, t, buf):
        """Called by the forward for every t; keeps what the top level needs.
        In store-all mode the forward hands over buffers it will not reuse."""
        if self.store_all:
            self.top[t] = buf
        elif t % self.stride == 0:
            self.top[t] = self.lat.clone_state(buf)

    def _recompute(self, p_start, t0, t1, keep_every):
        """{t: P_t} for t = t0 and every t0 < t < t1 with t % keep_every == 0, from P_t0."""
        lat = self.lat
        if keep_every == 1:
            bufs = [p_start] + lat.new_states(t1 - 1 - t0)
            lat.chain(bufs, step=t0)
            return {t0 + k: b for k, b in enumerate(bufs)}
        out = {t0: p_start}
        sa, sb = lat.new_state(), lat.new_state()
        cur, t = p_start, t0
        while t < t1 - 1:
            end = min(t1 - 1, (t // keep_every + 1) * keep_every)
            cur = lat.pingpong(cur, sa, sb, end - t, step=t)
            if end % keep_every == 0:
                out[end] = lat.clone_state(cur)
            t = end
        return out

    def segments(self):
        """(t0, t1, {t: P_t for t0 <= t < t1}) for consecutive segments, last first."""
        yield from self._seg(self.top, 0, self.N, 0)

    def _seg(self, have, t0, t1, level):
        if self.levels[level] == 1:
            yield t0, t1, have
            if level > 0:
                # a caller's loop variable still names this dict while the
                # next segment is recomputed: release its states here, or
                # two segments' worth would be alive at once
                have.clear()
            return
        starts = sorted(t for t in have if t0 <= t < t1)
        ends = starts[1:] + [t1]
        for s, e in reversed(list(zip(starts, ends))):
            sub = self._recompute(have[s], s, e, self.levels[level + 1])
            yield from self._seg(sub, s, e, level + 1)
            del sub

    def descending(self):
        """Generator of (t, P_t) for t from N-1 down to 0."""
        for t0, t1, have in self.segments():
            for t in range(t1 - 1, t0 - 1, -1):
                yield t, have[t]


class _RolloutFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, q, porosity, u_in, cfg):
        # outputs the loss does not use arrive as None instead of zero tensors
        # (no full-grid seed reads or adds for them)
        ctx.set_materialize_grads(False)
        n_steps = cfg["n_steps"]
        vel = [float(v) for v in u_in.detach().reshape(-1).to("cpu", torch.float64).tolist()]
        lat = Lattice(cfg["nx"], cfg["ny"], cfg["nz"], tau=cfg["tau"], inlet_velocity=vel, rho_out=cfg["rho_out"],
                      collision=cfg["collision"], smagorinsky_cs=cfg["smagorinsky_cs"], solid=cfg["solid"],
                      porosity=porosity.detach() if porosity is not None else None,
                      q=q.detach() if q is not None else None, guard=cfg["guard"],
                      guard_velocity=cfg["guard_velocity"], periodic_yz=True, storage=cfg["storage"],
                      offset=cfg["offset"], device=f0.device, layout=cfg["layout"],
                      cells_per_thread=cfg["cells_per_thread"], rounding=cfg["rounding"], seed=cfg["seed"])
        ck = _Checkpoints(lat, cfg["checkpoint"], n_steps)
        s0 = lat.to_state(f0.detach())
        a = lat.new_state()
        fpart = torch.zeros((n_steps, max(lat.slots, 1), 3), dtype=torch.float64, device=lat.device)
        lat.collide(s0, a)
        fp = fpart if lat.slots else None
        if ck.store_all:
            # every P_t in a buffer of its own, all carved from one allocation
            bufs = [a] + lat.new_states(n_steps - 1)
            lat.chain(bufs, fp, step=0)
            for t, buf in enumerate(bufs):
                ck.record(t, buf)
            cur = bufs[-1]
        else:
            b = lat.new_state()
            cur, t = a, 0
            while t < n_steps - 1:
                ck.record(t, cur)
                end = min(n_steps - 1, (t // ck.stride + 1) * ck.stride)
                cur = lat.pingpong(cur, a, b, end - t, None if fp is None else fp[t:end], step=t)
                t = end
            ck.record(n_steps - 1, cur)
        s_final = lat.new_state()
        lat.stream(cur, s_final, fpart[n_steps - 1] if lat.slots else None, step=n_steps - 1)
        forces = lat.reduce_forces(fpart, n_steps) if lat.slots else torch.zeros(
            (n_steps, 3), dtype=torch.float64, device=lat.device)
        p_last = lat.clone_state(cur) if lat.storage in ("f16", "bf16") else cur
        ctx.lat, ctx.ck, ctx.s0 = lat, ck, s0
        ctx.cfg = cfg
        ctx.has_q = q is not None
        ctx.has_phi = porosity is not None
        ctx.f0_dtype = f0.dtype
        ctx.q_dtype = q.dtype if q is not None else None
        ctx.phi_dtype = porosity.dtype if porosity is not None else None
        ctx.u_dtype, ctx.u_device = u_in.dtype, u_in.device
        s_out = lat.from_state(s_final)
        p_out = lat.from_state(p_last)
        if lat.storage in ("f32", "f64"):
            s_out = s_out.clone()
            p_out = p_out.clone()
        return s_out.to(f0.dtype), p_out.to(f0.dtype), forces

    @staticmethod
    def backward(ctx, gS, gP, gF):
        lat, ck = ctx.lat, ctx.ck
        n_steps = ctx.cfg["n_steps"]
        real = torch.float64 if lat.storage == "f64" else torch.float32
        dev = lat.device
        nx, ny, nz, N = lat.nx, lat.ny, lat.nz, lat.N
        n = lat.links.n
        want_q = ctx.needs_input_grad[1] and ctx.has_q and n > 0
        want_phi = ctx.needs_input_grad[2] and ctx.has_phi
        want_u = ctx.needs_input_grad[3]
        want_f0 = ctx.needs_input_grad[0]

        ad = AdjointDesc()
        keep = []
        # link-cell cotangents of two consecutive steps: the chain folds each
        # step's link additions into the next step's loads (fold_links), so no
        # scatter plan and no extra launch per step
        mu = torch.zeros((19, max(n, 1)), dtype=real, device=dev)
        keep.append(mu)
        ad.mu_link = L.ptr(mu)
        fold = ctx.cfg["link_additions"] == "fold"
        if fold:
            mu_alt = torch.zeros_like(mu)
            keep.append(mu_alt)
            ad.mu_link_alt = L.ptr(mu_alt)
            ad.fold_links = 1
        else:
            # one scatter launch per step from a per-entry plan
            plan = scatter_plan(lat)
            if plan is not None:
                keep.extend(plan)
                ad.n_targets = int(plan[0].numel())
                ad.tgt_index, ad.tgt_start, ad.src, ad.coef = (L.ptr(t) for t in plan)
        dq = None
        if want_q:
            dq = torch.zeros((18, n), dtype=torch.float64, device=dev)
            ad.dq_acc = L.ptr(dq)
        phi_acc = None
        if want_phi:
            phi_acc = torch.zeros(N, dtype=torch.float64, device=dev)
            ad.phi_acc = L.ptr(phi_acc)
        u_acc = None
        if want_u:
            u_acc = torch.zeros((ny * nz, 3), dtype=torch.float64, device=dev)
            ad.u_acc = L.ptr(u_acc)
        if gF is None:
            gF = torch.zeros((n_steps, 3), dtype=torch.float64, device=dev)
        gF = gF.detach().to(torch.float64).contiguous()
        stream = L.stream_handle(dev)
        dptr = ctypes.byref(lat.desc)
        aptr = ctypes.byref(ad)

        lam_dtype, scale = real, 1.0
        if ctx.cfg["adjoint_storage"] == "f16":
            # lambda stored as IEEE half times a power of two that puts the
            # largest seed at 2^12 (16x headroom below the half maximum)
            seed = 3.0 * float(gF.abs().max()) if gF.numel() else 0.0
            for g in (gS, gP):
                if g is not None:
                    seed = max(seed, float(g.detach().abs().max()))
            scale = 2.0 ** (12 - math.ceil(math.log2(seed))) if seed > 0.0 else 1.0
            lam_dtype = torch.float16
            ad.lam_storage = L.LBM_F16DEV
            ad.lam_scale = scale
        lam_a = lat.new_state(lam_dtype)
        lam_b = lat.new_state(lam_dtype)
        g_in = lat.new_state(lam_dtype)
        if gS is not None:
            g_in.copy_(gS.detach().permute(0, 3, 2, 1).to(real) * scale)
        else:
            g_in.zero_()
        cur, nxt = lam_a, lam_b
        first = True
        for t0, t1, have in ck.segments():
            if first:
                # epilogue transpose at t = N-1: lam = dL/dS_N -> dL/dP_{N-1}
                t = n_steps - 1
                assert t1 == n_steps
                L.call("lbm_step_adjoint", dptr, aptr, L.LBM_MODE_STREAM, L.ptr(have[t]), L.ptr(g_in), L.ptr(cur),
                       L.ptr(gF[t]), stream)
                if fold:
                    L.call("lbm_adjoint_link_finish", dptr, aptr, L.ptr(mu), L.ptr(cur), stream)
                else:
                    L.call("lbm_adjoint_scatter", dptr, aptr, L.ptr(cur), L.ptr(gF[t]), stream)
                if gP is not None:
                    cur.copy_(cur.to(real) + gP.detach().permute(0, 3, 2, 1).to(real) * scale)
                t1 -= 1
                first = False
            n = t1 - t0
            if n > 0:
                L.call("lbm_adjoint_chain", dptr, aptr, n, L.ptr_array([have[t] for t in range(t0, t1)]),
                       L.ptr(cur), L.ptr(nxt), L.ptr(gF[t0]), stream)
                if n % 2 == 1:
                    cur, nxt = nxt, cur
        if lam_dtype == torch.float16 and not bool(torch.isfinite(cur).all()):
            raise FloatingPointError("adjoint cotangents left the float16 range; use adjoint_storage='f32'")
        # cur = dL/dP_0 -> prologue transpose (output in the arithmetic type)
        g_f0 = None
        if want_f0 or want_phi:
            g_s0 = nxt if lam_dtype == real else lat.new_state(real)
            L.call("lbm_step_adjoint", dptr, aptr, L.LBM_MODE_COLLIDE, L.ptr(ctx.s0), L.ptr(cur), L.ptr(g_s0),
                   None, stream)
            if want_f0:
                g_f0 = F.pop_view(g_s0).to(ctx.f0_dtype).clone()
        g_q = None
        if want_q:
            dense = torch.zeros((19, N), dtype=ctx.q_dtype, device=dev)
            cells = lat.links.cell.long()
            dense[1:, :].index_copy_(1, cells, dq.to(ctx.q_dtype))
            g_q = F.pop_view(dense.reshape(19, nz, ny, nx))
        g_phi = F.scalar_view(phi_acc.reshape(nz, ny, nx)).to(ctx.phi_dtype) if want_phi else None
        # on the velocity's own device: a host tensor with requires_grad is accepted
        g_u = u_acc.sum(0).to(device=ctx.u_device, dtype=ctx.u_dtype) if want_u else None
        return g_f0, g_q, g_phi, g_u, None


def rollout(f0: torch.Tensor, n_steps: int, *, tau, q: Optional[torch.Tensor] = None,
            porosity: Optional[torch.Tensor] = None, solid: Optional[torch.Tensor] = None,
            inlet_velocity: Union[Sequence[float], torch.Tensor] = (0.05, 0.0, 0.0), rho_out: float = 1.0,
            collision: str = "bgk", smagorinsky_cs: Optional[float] = None, guard: str = "none",
            guard_velocity=None, storage: str = "f32", offset="weights",
            checkpoint: Union[str, int, Sequence[int]] = "all", adjoint_storage: Optional[str] = None,
            layout: str = "rows", cells_per_thread: int = 0, link_additions: str = "fold",
            rounding: str = "stochastic", seed: int = 0):
    """Differentiable fused rollout: (S_N, P_{N-1}, forces[N, 3] float64).

    Step order as in the reference's drivers: [pre guard] -> [Smagorinsky]
    -> BGK|MRT -> soft bounce (porosity) or hard bounce (solid >= 1) -> stream
    with Bouzidi (q) -> Zou-He inlet/outlet -> [post guard].
    ``adjoint_storage`` "f32" keeps the cotangents in the arithmetic type,
    "f16" stores them as scaled IEEE half (half storages only; default for
    them), halving the adjoint's cotangent traffic.  ``layout`` places the
    state and cotangent buffers ("rows" or "apart", as for ``Lattice``);
    ``cells_per_thread`` selects the sweeps' addressing as for ``Lattice``.
    ``link_additions`` "fold" (default) lets each adjoint step add the
    previous step's Bouzidi and force cotangents as it loads them (one launch
    per step); "scatter" adds them with a separate launch per step from a
    per-entry plan (the same sums in another order).  ``rounding`` and
    ``seed`` set the 16-bit formats' rounding as for ``Lattice``; the
    checkpoint recompute draws the same dither, so it reproduces the forward.
    """
    nx, ny, nz = F.dims_of_pop(f0)
    if isinstance(inlet_velocity, torch.Tensor):
        u = inlet_velocity.reshape(3)
    else:
        u = torch.tensor([float(v) for v in inlet_velocity], dtype=torch.float64, device=f0.device)
    if adjoint_storage is None:
        adjoint_storage = "f16" if storage in ("f16", "bf16") else "f32"
    if adjoint_storage not in ("f32", "f16") or (adjoint_storage == "f16" and storage == "f64"):
        raise ValueError("adjoint_storage must be 'f32', or 'f16' with float arithmetic")
    cfg = dict(n_steps=int(n_steps), nx=nx, ny=ny, nz=nz, tau=F.scalar_float(tau), rho_out=float(rho_out),
               collision=collision, smagorinsky_cs=smagorinsky_cs, solid=solid, guard=guard,
               guard_velocity=guard_velocity, storage=storage, offset=offset, checkpoint=checkpoint,
               adjoint_storage=adjoint_storage, layout=layout, cells_per_thread=int(cells_per_thread),
               link_additions=link_additions, rounding=rounding, seed=int(seed))
    if link_additions not in ("fold", "scatter"):
        raise ValueError("link_additions must be 'fold' or 'scatter'")
    if rounding not in ("nearest", "stochastic"):
        raise ValueError("rounding must be 'nearest' or 'stochastic'")
    if int(n_steps) < 1:
        raise ValueError("n_steps must be >= 1")
    return _RolloutFn.apply(f0, q, porosity, u, cfg)


def drag_series(forces: torch.Tensor, inlet_velocity: float, area, rho: float = 1.0) -> torch.Tensor:
    """Cd_t = |F_t,x| / (0.5 rho U^2 A + 1e-10) per step, differentiable in the
    forces (jax_lbm/core/forces.py:237-276 applied to each row)."""
    return forces[:, 0].abs() / (0.5 * rho * inlet_velocity * inlet_velocity * area + 1e-10)
