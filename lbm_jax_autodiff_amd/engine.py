"""Fused rollout engine: the hot path of jax_lbm on liblbm_hip.so.

A ``Lattice`` compiles one geometry and configuration (collision model,
bounce mode, guard variant, storage precision) into the ``lbm_step_desc`` of
the C ABI and runs timesteps with one kernel launch each.  The carried state
is the post-collision population P_t; ``run`` turns an initial state S_0 into
(S_N, P_{N-1}, forces[N, 3]) exactly as the reference's scan loops return
(f_final, f_post) and compute momentum_exchange(f_post, f_final, q) every step
(examples/10_cd_convergence.py:69-97, examples/11_optimisation_v2.py:141-167,
jax_lbm/state.py:87-230).
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _field as F
from . import _lib as L

_STORAGE = {
    "f32": (L.LBM_F32, torch.float32, torch.float32),
    "f64": (L.LBM_F64, torch.float64, torch.float64),
    "f16": (L.LBM_F16DEV, torch.float16, torch.float32),
    "bf16": (L.LBM_BF16DEV, torch.bfloat16, torch.float32),
}
_ROUNDING = {"nearest": L.LBM_ROUND_NEAREST, "stochastic": L.LBM_ROUND_STOCHASTIC}
_GUARD = {"none": L.LBM_GUARD_NONE, "post": L.LBM_GUARD_POST, "pre": L.LBM_GUARD_PRE}
_W = [1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12
_E = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 1, 0),
      (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1),
      (0, -1, -1)]


def _f32(x: float) -> float:
    return float(torch.tensor(x, dtype=torch.float32).item())


def equilibrium_values(rho: float, u: Sequence[float]):
    """The 19 values of f_eq(rho, u) in float64 (host; used for offsets)."""
    ux, uy, uz = (float(v) for v in u)
    usq = ux * ux + uy * uy + uz * uz
    out = []
    for w, (ex, ey, ez) in zip(_W, _E):
        eu = ex * ux + ey * uy + ez * uz
        out.append(_f32(w) * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq))
    return out


class Links:
    """Compiled Bouzidi links of a dense q: link cells, their 18 q values and flags."""

    def __init__(self, n, cell, q, dq, flags, dense_q):
        self.n = n
        self.cell = cell          # (n,) int32 cell index (lattice storage)
        self.q = q                # (18, n) float64
        self.dq = dq              # (18, n) float64 or None
        self.flags = flags        # (N,) uint8
        self.dense_q = dense_q    # lattice-storage q used to build them


def compile_links(q_storage: Optional[torch.Tensor], solid_storage: Optional[torch.Tensor], nx, ny, nz, device,
                  dqdr_storage: Optional[torch.Tensor] = None) -> Links:
    """Dense q (lattice storage, float32/float64) -> link list + per-cell flags."""
    N = nx * ny * nz
    if q_storage is None:
        flags = torch.zeros(N, dtype=torch.uint8, device=device)
        if solid_storage is not None:
            flags = (solid_storage.reshape(-1) >= 1).to(torch.uint8)
        return Links(0, None, None, None, flags, None)
    tag = F.dtype_tag(q_storage.dtype)
    nb = (N + 255) // 256
    counts = torch.empty(nb, dtype=torch.int32, device=device)
    offsets = torch.empty(nb, dtype=torch.int32, device=device)
    total = torch.empty(1, dtype=torch.int32, device=device)
    stream = L.stream_handle(device)
    L.call("lbm_links_count", tag, L.ptr(q_storage), L.ptr(counts), L.ptr(offsets), L.ptr(total), nx, ny, nz, stream)
    n = int(total.item())
    cell = torch.empty(max(n, 1), dtype=torch.int32, device=device)
    lq = torch.empty((18, max(n, 1)), dtype=torch.float64, device=device)
    ldq = torch.empty((18, max(n, 1)), dtype=torch.float64, device=device) if dqdr_storage is not None else None
    flags = torch.empty(N, dtype=torch.uint8, device=device)
    if dqdr_storage is not None:
        dqdr_storage = dqdr_storage.to(q_storage.dtype).contiguous()
    L.call("lbm_links_build", tag, L.ptr(q_storage), L.ptr(dqdr_storage), L.ptr(solid_storage), L.ptr(offsets), n,
           L.ptr(cell), L.ptr(lq), L.ptr(ldq), L.ptr(flags), nx, ny, nz, stream)
    if n == 0:
        cell, lq, ldq = None, None, None
    return Links(n, cell, lq, ldq, flags, q_storage)


class Lattice:
    """One compiled configuration of the fused D3Q19 step.

    Parameters mirror LBMParams (jax_lbm/state.py:57-84) plus the variants the
    reference's drivers use: ``bounce`` hard (solid >= 1 reversal,
    state.py:146-149) or soft (porosity blend, smooth.py:151-155);
    ``guard`` none, post (state.py:158-169, reset to eq(1, 0)) or pre
    (examples/10:71-75, reset to eq(1, guard_velocity)).
    ``storage`` f32 / f64 / f16 / bf16 (half formats store f_i - offset_i
    with float32 arithmetic); ``offset`` "weights" (w_i) or "inlet"
    (f_eq(1, inlet_velocity)) or 19 numbers.  ``cells_per_thread`` 0
    (default, the fastest) runs one cell per thread and, where nx is a
    multiple of 64, keeps each wave's row addresses in scalar registers
    (16-bit populations and the adjoint); 1 forces per-lane addresses; 2 or 4
    (when it divides nx) update that many x-consecutive cells per thread with
    vector loads.
    ``layout`` places the state buffers of the fused path: "rows" (default)
    interleaves the populations of every x-row, [nz][ny][19][nx], so a row's
    19 populations are one contiguous block; "apart" keeps whole populations
    apart, [19][pop_stride] with cells x-fastest.  ``pop_stride`` is the
    element distance between consecutive populations (rows: nx by default;
    apart: the library's padded recommendation by default, N for the dense
    (19, nz, ny, nx) array).
    ``rounding`` of the 16-bit formats: "stochastic" (default) -- each
    stored deviation is rounded up or down with probabilities that make its
    expectation the exact value (gfx950's stochastic-rounding conversions,
    fed a hash of (``seed``, time index, cell, population), so recomputing a
    step reproduces its bits) -- or "nearest" (even).  Rounding to nearest
    leaves a steady flow's stored values with a fixed bias (a change smaller
    than half a unit in the last place never lands); stochastic rounding
    makes those errors average out over the steps.
    """

    def __init__(self, nx, ny, nz, *, tau, inlet_velocity=(0.05, 0.0, 0.0), rho_out=1.0, collision="bgk",
                 smagorinsky_cs=None, solid=None, porosity=None, q=None, guard="none", guard_velocity=None,
                 periodic_yz=True, storage="f32", offset="weights", device=None, dqdr=None,
                 cells_per_thread=0, pop_stride=None, layout="rows", rounding="stochastic", seed=0):
        if storage not in _STORAGE:
            raise ValueError(f"storage must be one of {sorted(_STORAGE)}")
        if guard not in _GUARD:
            raise ValueError(f"guard must be one of {sorted(_GUARD)}")
        if collision not in ("bgk", "mrt"):
            raise ValueError("collision must be 'bgk' or 'mrt'")
        if rounding not in _ROUNDING:
            raise ValueError(f"rounding must be one of {sorted(_ROUNDING)}")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.N = self.nx * self.ny * self.nz
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.storage = storage
        self.tag, self.buf_dtype, self.real = _STORAGE[storage]
        self.collision = collision
        self.guard = guard
        self.inlet_velocity = [float(v) for v in inlet_velocity]
        self.tau = F.scalar_float(tau)

        grid, pops = (self.nx, self.ny, self.nz), (19, self.nx, self.ny, self.nz)
        for t, name, shape in ((solid, "solid", grid), (porosity, "porosity", grid), (q, "q", pops),
                               (dqdr, "dqdr", pops)):
            if t is not None:
                F.require_shape(t, name, shape, self.device)
        solid_s = F.scalar_storage(solid.to(torch.int32)) if solid is not None else None
        self.solid_storage = solid_s
        self.porosity_storage = None
        if porosity is not None:
            self.porosity_storage = F.scalar_storage(porosity.detach().to(self.real))
        qs = None
        if q is not None:
            qs = F.pop_storage(q.detach())
            if qs.dtype not in (torch.float32, torch.float64):
                qs = qs.to(torch.float32)
        dq_s = F.pop_storage(dqdr.detach()) if dqdr is not None else None
        self.links = compile_links(qs, solid_s, self.nx, self.ny, self.nz, self.device, dq_s)

        d = L.StepDesc()
        d.nx, d.ny, d.nz = self.nx, self.ny, self.nz
        d.storage = self.tag
        d.collision = L.LBM_COLL_MRT if collision == "mrt" else L.LBM_COLL_BGK
        d.smagorinsky = 1 if smagorinsky_cs is not None else 0
        d.smag_cs = float(smagorinsky_cs) if smagorinsky_cs is not None else 0.1
        if porosity is not None:
            d.bounce = L.LBM_BOUNCE_SOFT
        elif solid is not None:
            d.bounce = L.LBM_BOUNCE_HARD
        else:
            d.bounce = L.LBM_BOUNCE_NONE
        d.guard = _GUARD[guard]
        d.periodic_yz = 1 if periodic_yz else 0
        d.n_links = self.links.n
        d.tau = self.tau
        for k in range(3):
            d.u_in[k] = self.inlet_velocity[k]
        d.rho_out = float(rho_out)
        gv = self.inlet_velocity if guard_velocity is None else [float(v) for v in guard_velocity]
        for k in range(3):
            d.u_ref[k] = gv[k]
        if isinstance(offset, str):
            if offset == "weights":
                offs = [_f32(w) for w in _W]
            elif offset == "inlet":
                offs = equilibrium_values(1.0, self.inlet_velocity)
            else:
                raise ValueError("offset must be 'weights', 'inlet' or 19 numbers")
        else:
            offs = [float(v) for v in offset]
        if storage in ("f16", "bf16"):
            offs = [_f32(v) for v in offs]
        else:
            offs = [0.0] * 19
        for i in range(19):
            d.offset[i] = offs[i]
        d.flags = L.ptr(self.links.flags)
        d.porosity = L.ptr(self.porosity_storage)
        d.link_cell = L.ptr(self.links.cell)
        d.link_q = L.ptr(self.links.q)
        d.cells_per_thread = int(cells_per_thread)
        d.rounding = _ROUNDING[rounding]
        d.seed = int(seed) & 0xFFFFFFFF
        d.step = 0
        self.rounding = rounding
        lib = L.load()
        if layout == "rows":
            self.P = self.nx if pop_stride is None else int(pop_stride)
            self.RS = 19 * self.P
            ok = self.P >= self.nx
        elif layout == "apart":
            self.P = int(lib.lbm_state_stride(ctypes.byref(d))) if pop_stride is None else int(pop_stride)
            self.RS = self.nx
            ok = self.P >= self.N
        else:
            raise ValueError("layout must be 'rows' or 'apart'")
        self.layout = layout
        self.extent = 18 * self.P + (self.ny * self.nz - 1) * self.RS + self.nx
        if not ok or self.extent >= 2 ** 31:
            raise ValueError(f"pop_stride {self.P} does not fit layout {layout!r} for grid {nx}x{ny}x{nz}")
        d.pop_stride = self.P
        d.row_stride = self.RS
        self.desc = d
        self.slots = int(lib.lbm_step_force_slots(ctypes.byref(d)))

    # ---- state buffers ---------------------------------------------------
    def new_state(self, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
        """Uninitialised buffer in the state layout, viewed as (19, nz, ny, nx)
        (storage dtype by default; cotangent buffers pass their real dtype)."""
        flat = torch.empty(self.extent, dtype=dtype or self.buf_dtype, device=self.device)
        return flat.as_strided((19, self.nz, self.ny, self.nx), (self.P, self.ny * self.RS, self.RS, 1))

    def new_states(self, n: int, dtype: Optional[torch.dtype] = None) -> list:
        """``n`` uninitialised state buffers carved from one allocation, each
        starting on a 256-byte boundary (one allocator call instead of n)."""
        if n <= 0:
            return []
        dt = dtype or self.buf_dtype
        esize = torch.empty((), dtype=dt).element_size()
        step = -(-self.extent * esize // 256) * 256 // esize
        flat = torch.empty(n * step, dtype=dt, device=self.device)
        shape = (19, self.nz, self.ny, self.nx)
        strides = (self.P, self.ny * self.RS, self.RS, 1)
        return [flat.as_strided(shape, strides, k * step) for k in range(n)]

    def clone_state(self, buf: torch.Tensor) -> torch.Tensor:
        out = self.new_state(buf.dtype)
        out.copy_(buf)
        return out

    def to_state(self, f: torch.Tensor, step: int = 0) -> torch.Tensor:
        """(19, nx, ny, nz) float tensor -> state buffer in the storage format
        (16-bit formats: rounded as ``rounding`` says, with the dither of time
        index ``step``)."""
        F.require_cuda(f, "f")
        fs = f.permute(0, 3, 2, 1)
        if self.storage in ("f32", "f64"):
            out = self.new_state()
            out.copy_(fs)
            return out
        src = self.new_state(torch.float32)
        src.copy_(fs)
        out = self.new_state()
        self.desc.step = int(step)
        L.call("lbm_convert_state", ctypes.byref(self.desc), L.LBM_F32, L.ptr(src), self.tag, L.ptr(out),
               L.stream_handle(self.device))
        return out

    def from_state(self, buf: torch.Tensor) -> torch.Tensor:
        """State buffer -> float32/float64 tensor indexed (19, nx, ny, nz)."""
        if self.storage in ("f32", "f64"):
            return F.pop_view(buf)
        out = self.new_state(torch.float32)
        L.call("lbm_convert_state", ctypes.byref(self.desc), self.tag, L.ptr(buf), L.LBM_F32, L.ptr(out),
               L.stream_handle(self.device))
        return F.pop_view(out)

    # ---- single launches -------------------------------------------------
    # `step` is the time index t of the launch (P_{t+1} from P_t; the
    # prologue and the epilogue use their own t): with stochastic rounding it
    # selects the dither, so a recomputed step reproduces its stored values.
    def launch(self, mode: int, src: torch.Tensor, dst: torch.Tensor, fpart: Optional[torch.Tensor] = None,
               step: int = 0):
        self.desc.step = int(step)
        L.call("lbm_step", ctypes.byref(self.desc), int(mode), L.ptr(src), L.ptr(dst), L.ptr(fpart),
               L.stream_handle(self.device))

    def collide(self, s_buf, p_buf, step: int = 0):
        """P_t = bounce(collide(S_t)) (rollout prologue)."""
        self.launch(L.LBM_MODE_COLLIDE, s_buf, p_buf, step=step)

    def step(self, p_in, p_out, fpart=None, step: int = 0):
        """P_{t+1} from P_t in one launch; fpart receives this step's force partials."""
        self.launch(L.LBM_MODE_FULL, p_in, p_out, fpart, step=step)

    def stream(self, p_in, s_out, fpart=None, step: int = 0):
        """S_{t+1} from P_t (rollout epilogue)."""
        self.launch(L.LBM_MODE_STREAM, p_in, s_out, fpart, step=step)

    def chain(self, states, fpart: Optional[torch.Tensor] = None, step: int = 0):
        """len(states) - 1 full steps in one host call: states[k] -> states[k+1]
        (time indices step, step + 1, ...).  fpart, if given, is a contiguous
        (n, slots, 3) float64 block receiving each step's force partials."""
        n = len(states) - 1
        if n <= 0:
            return
        fp, stride = 0, 0
        if fpart is not None and self.slots > 0:
            assert fpart.is_contiguous() and fpart.shape[0] >= n
            fp, stride = fpart.data_ptr(), 3 * self.slots
        self.desc.step = int(step)
        L.call("lbm_step_chain", ctypes.byref(self.desc), n, L.ptr_array(states), fp, stride,
               L.stream_handle(self.device))

    def chain_forces_noalloc(self, states, force: torch.Tensor, step: int = 0):
        """len(states) - 1 full steps and their forces (float64, (n, 3)) in one
        call that allocates no device memory: the force partials live in the
        library's own device array (lbm_step_chain_forces_noalloc), which
        bounds n * slots * 3 by LBM_NOALLOC_PARTIALS and is shared by every
        call on the device."""
        n = len(states) - 1
        if n <= 0:
            raise ValueError("need at least two states")
        if n * self.slots * 3 > L.LBM_NOALLOC_PARTIALS:
            raise ValueError(f"{n} steps x {self.slots} force slots exceed the no-allocation partials array "
                             f"({L.LBM_NOALLOC_PARTIALS} doubles)")
        F.require_shape(force, "force", (n, 3), self.device)
        if force.dtype != torch.float64 or not force.is_contiguous():
            raise ValueError("force must be a contiguous float64 tensor")
        self.desc.step = int(step)
        L.call("lbm_step_chain_forces_noalloc", ctypes.byref(self.desc), n, L.ptr_array(states), L.ptr(force),
               L.stream_handle(self.device))

    def pingpong(self, cur, spare_a, spare_b, n, fpart=None, step: int = 0):
        """n full steps from cur (time index `step`) through two spare buffers
        (never writing cur unless it is one of them); returns the buffer
        holding the result."""
        if n <= 0:
            return cur
        first = spare_a if cur is not spare_a else spare_b
        second = spare_b if first is spare_a else spare_a
        seq = [cur] + [first if k % 2 == 0 else second for k in range(n)]
        self.chain(seq, fpart, step=step)
        return seq[-1]

    def reduce_forces(self, fpart: torch.Tensor, T: int) -> torch.Tensor:
        force = torch.empty((T, 3), dtype=torch.float64, device=self.device)
        if T > 0:
            L.call("lbm_force_reduce", L.ptr(fpart), int(T), int(self.slots), L.ptr(force),
                   L.stream_handle(self.device))
        return force

    # ---- rollouts --------------------------------------------------------
    def run(self, f0: torch.Tensor, n_steps: int, *, forces: bool = True, keep_every: int = 0,
            keep: Optional[list] = None):
        """Run n_steps timesteps from S_0 = f0.

        Returns (S_N, P_{N-1}, F) with F[t] = momentum_exchange(P_t, S_{t+1})
        in float64 (None if forces=False).  When ``keep`` is a list, the P_t
        buffers with t % keep_every == 0 are appended to it (checkpoints).
        """
        n_steps = int(n_steps)
        if n_steps < 1:
            raise ValueError("n_steps must be >= 1")
        s0 = self.to_state(f0)
        a = self.new_state()
        b = self.new_state()
        fpart = None
        if forces and self.slots > 0:
            fpart = torch.zeros((n_steps, self.slots, 3), dtype=torch.float64, device=self.device)
        self.collide(s0, a)
        cur = a
        t = 0
        while t < n_steps - 1:
            if keep is not None and keep_every and t % keep_every == 0:
                keep.append((t, self.clone_state(cur)))
            end = n_steps - 1
            if keep is not None and keep_every:
                end = min(end, (t // keep_every + 1) * keep_every)
            cur = self.pingpong(cur, a, b, end - t, None if fpart is None else fpart[t:end], step=t)
            t = end
        if keep is not None and keep_every and (n_steps - 1) % keep_every == 0:
            keep.append((n_steps - 1, self.clone_state(cur)))
        s_final = s0  # reuse the prologue input buffer for S_N
        self.stream(cur, s_final, None if fpart is None else fpart[n_steps - 1], step=n_steps - 1)
        F_hist = None
        if forces:
            F_hist = self.reduce_forces(fpart, n_steps) if fpart is not None else torch.zeros(
                (n_steps, 3), dtype=torch.float64, device=self.device)
        return self.from_state(s_final), self.from_state(cur), F_hist
