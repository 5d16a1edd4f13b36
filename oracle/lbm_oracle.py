"""CPU oracle for the jax_lbm per-timestep hot path (TEST INFRASTRUCTURE ONLY).

This module is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``lbm_jax_autodiff_amd``) runs on the HIP
library and fails loudly when that library is missing; it never calls into
this file.

It restates, in PyTorch on the CPU (float32 or float64, differentiable with
``torch.autograd``), the algorithm of the reference package ``jax_lbm``
(MarcosAsh/LBM_JAX_Autodiff).  Every function cites the reference file:line it
follows.  The layout is the reference's: populations ``f[i, x, y, z]`` of shape
``(19, nx, ny, nz)``.

How each part is pinned (tests in ``tests/test_oracle.py``; the published
values are read from ``tests/golden/kats.json``):

* lattice tables, MRT basis     closed forms (weight moments, M M^T = diag(norms)),
                                 the reference's row-norm table
* equilibrium / moments          closed forms (rho, rho u, second moment), the
                                 FP64 KAT d(sum f_eq)/d rho = 64
* bgk / mrt / smagorinsky        fixed points, conservation, m_eq = M f_eq,
                                 tau_eff >= tau, planted-mistake tests
* stream (+Bouzidi)              q=1/2 identity, branch continuity, shift tests,
                                 the Poiseuille KAT (6 decimals, 60x20x4, 3000 steps)
* zou_he_inlet / zou_he_outlet   closed-form post-conditions, FP64 KAT
                                 d mean(rho)/dU = 0.611238
* momentum_exchange / Cd         no-link zero, Cd = 0.08 KAT
* smooth geometry, soft bounce   FP64 KAT dCd/dr = 93.33275 (exercises q(r) and
                                 porosity(r)), closed-form dq/dr vs finite differences
* projected_area                 box = 16, ground excluded

The reference itself cannot be imported (jax/jaxlib are absent), so FP32
bit-level agreement with XLA is unpinned; agreement is pinned at the values the
reference publishes.
"""

from __future__ import annotations

import functools
import math
from typing import NamedTuple, Optional, Sequence

import torch

# ---------------------------------------------------------------------------
# Lattice tables (jax_lbm/core/lattice.py:29-129).  The velocity ordering,
# weights, opposite map and MRT rate table must be identical to the
# reference's: they define the data layout and the collision model.
# ---------------------------------------------------------------------------

Q = 19
VELOCITIES = (
    (0, 0, 0),
    (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
    (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0),
    (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
    (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1),
)


def _opposite_table() -> tuple:
    out = []
    for e in VELOCITIES:
        neg = (-e[0], -e[1], -e[2])
        out.append(VELOCITIES.index(neg))
    return tuple(out)


OPPOSITE = _opposite_table()

# The reference keeps its weights and MRT rates as float32 arrays even when JAX
# runs in x64 (lattice.py:53-60, 106-126); its FP64 results therefore use the
# float32-rounded values.  ``faithful32=True`` reproduces that.
_W_EXACT = tuple(1.0 / 3.0 if i == 0 else (1.0 / 18.0 if i <= 6 else 1.0 / 36.0)
                 for i in range(Q))
CS2 = 1.0 / 3.0

# Fixed MRT rates (lattice.py:106-126); the stress slots 9, 11, 13, 14, 15 hold
# omega = 1/tau at run time (collision.py:162, lattice.py:129).
_MRT_RATES_FIXED = (0.0, 1.19, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 1.2,
                    0.0, 1.4, 0.0, 1.4, 0.0, 0.0, 0.0, 1.98, 1.98, 1.98)
MRT_STRESS = (9, 11, 13, 14, 15)
# Row norms ||M_k||^2 as tabulated by the reference (lattice.py:88-94).
MRT_ROW_NORMS_REF = (19.0, 2394.0, 252.0, 10.0, 40.0, 10.0, 40.0, 10.0, 40.0,
                     36.0, 72.0, 12.0, 24.0, 4.0, 4.0, 4.0, 8.0, 8.0, 8.0)


def _f32(x: float) -> float:
    return float(torch.tensor(x, dtype=torch.float32).item())


def weights(dtype=torch.float64, faithful32: bool = True) -> torch.Tensor:
    vals = [_f32(w) if faithful32 else w for w in _W_EXACT]
    return torch.tensor(vals, dtype=dtype)


def velocities(dtype=torch.float64) -> torch.Tensor:
    return torch.tensor(VELOCITIES, dtype=dtype)


def mrt_rates_fixed(dtype=torch.float64, faithful32: bool = True) -> torch.Tensor:
    vals = [_f32(s) if faithful32 else s for s in _MRT_RATES_FIXED]
    return torch.tensor(vals, dtype=dtype)


@functools.lru_cache(maxsize=None)
def mrt_matrix(dtype=torch.float64) -> torch.Tensor:
    """The d'Humieres D3Q19 moment matrix, built from its defining polynomials.

    Row order follows lattice.py:140-145 (rho, e, eps, jx, qx, jy, qy, jz, qz,
    3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz).  Built here from the
    polynomials of (cx, cy, cz) rather than from the reference's unrolled sums
    (lattice.py:184-209); tests check it against the reference's row-norm table.
    Cached per dtype: callers only read it.
    """
    rows = []
    for k in range(Q):
        row = []
        for (cx, cy, cz) in VELOCITIES:
            c2 = cx * cx + cy * cy + cz * cz
            val = {
                0: 1,
                1: 19 * c2 - 30,
                2: (21 * c2 * c2 - 53 * c2 + 24) // 2,
                3: cx, 4: (5 * c2 - 9) * cx,
                5: cy, 6: (5 * c2 - 9) * cy,
                7: cz, 8: (5 * c2 - 9) * cz,
                9: 3 * cx * cx - c2,
                10: (3 * c2 - 5) * (3 * cx * cx - c2),
                11: cy * cy - cz * cz,
                12: (3 * c2 - 5) * (cy * cy - cz * cz),
                13: cx * cy, 14: cy * cz, 15: cx * cz,
                16: (cy * cy - cz * cz) * cx,
                17: (cz * cz - cx * cx) * cy,
                18: (cx * cx - cy * cy) * cz,
            }[k]
            row.append(val)
        rows.append(row)
    return torch.tensor(rows, dtype=dtype)


def mrt_forward(f_cell: torch.Tensor) -> torch.Tensor:
    """m = M f for one cell or a field (leading axis 19).  lattice.py:132-211."""
    M = mrt_matrix(f_cell.dtype)
    return torch.einsum("ki,i...->k...", M, f_cell)


def mrt_inverse(m: torch.Tensor) -> torch.Tensor:
    """f = M^T diag(1/||M_k||^2) m.  lattice.py:214-339."""
    M = mrt_matrix(m.dtype)
    norms = torch.tensor(MRT_ROW_NORMS_REF, dtype=m.dtype)
    shape = (Q,) + (1,) * (m.dim() - 1)
    return torch.einsum("ki,k...->i...", M, m / norms.reshape(shape))


def mrt_equilibrium_moments(rho: torch.Tensor, u: torch.Tensor) -> torch.Tensor:
    """Equilibrium moments (lattice.py:342-391); u has leading axis 3."""
    ux, uy, uz = u[0], u[1], u[2]
    u2 = ux * ux + uy * uy + uz * uz
    z = torch.zeros_like(rho)
    dxx = 2.0 * ux * ux - uy * uy - uz * uz
    dww = uy * uy - uz * uz
    return torch.stack([
        rho,
        -11.0 * rho + 19.0 * rho * u2,
        3.0 * rho - 5.5 * rho * u2,
        rho * ux, -2.0 / 3.0 * rho * ux,
        rho * uy, -2.0 / 3.0 * rho * uy,
        rho * uz, -2.0 / 3.0 * rho * uz,
        rho * dxx, -0.5 * rho * dxx,
        rho * dww, -0.5 * rho * dww,
        rho * ux * uy, rho * uy * uz, rho * ux * uz,
        z, z, z,
    ])


# ---------------------------------------------------------------------------
# Equilibrium and moments (jax_lbm/core/equilibrium.py)
# ---------------------------------------------------------------------------

def _bshape(t: torch.Tensor, nd: int) -> torch.Tensor:
    return t.reshape((Q,) + (1,) * nd)


def compute_density(f: torch.Tensor) -> torch.Tensor:
    """rho = sum_i f_i.  equilibrium.py:22-48."""
    return f.sum(dim=0)


def compute_velocity(f: torch.Tensor, rho: torch.Tensor) -> torch.Tensor:
    """u = sum_i f_i e_i / rho (no zero guard).  equilibrium.py:51-87."""
    e = velocities(f.dtype)
    return torch.einsum("i...,ia->a...", f, e) / rho.unsqueeze(0)


def equilibrium(rho: torch.Tensor, u: torch.Tensor, faithful32: bool = True) -> torch.Tensor:
    """Second-order D3Q19 equilibrium.  equilibrium.py:90-151."""
    dt = torch.promote_types(rho.dtype, u.dtype)
    e = velocities(dt)
    eu = torch.einsum("ia,a...->i...", e, u)
    usq = (u * u).sum(dim=0)
    w = _bshape(weights(dt, faithful32), rho.dim())
    return w * rho.unsqueeze(0) * (1.0 + eu / CS2 + (eu * eu) / (2.0 * CS2 * CS2)
                                   - usq.unsqueeze(0) / (2.0 * CS2))


# ---------------------------------------------------------------------------
# Collision (jax_lbm/core/collision.py)
# ---------------------------------------------------------------------------

def _omega_field(tau, like: torch.Tensor):
    if isinstance(tau, torch.Tensor) and tau.dim() > 0:
        return (1.0 / tau).unsqueeze(0)
    return 1.0 / tau


def bgk(f: torch.Tensor, tau, faithful32: bool = True) -> torch.Tensor:
    """f - (f - f_eq(rho(f), u(f)))/tau; tau scalar or per cell.  collision.py:46-100."""
    rho = compute_density(f)
    u = compute_velocity(f, rho)
    feq = equilibrium(rho, u, faithful32)
    return f - _omega_field(tau, f) * (f - feq)


def mrt(f: torch.Tensor, tau, faithful32: bool = True) -> torch.Tensor:
    """Moment-space relaxation with omega in the stress slots.  collision.py:103-184."""
    rho = compute_density(f)
    u = compute_velocity(f, rho)
    m = mrt_forward(f)
    meq = mrt_equilibrium_moments(rho, u)
    nd = f.dim() - 1
    s = mrt_rates_fixed(f.dtype, faithful32)
    s = s.reshape((Q,) + (1,) * nd).expand((Q,) + tuple(f.shape[1:])).clone()
    if isinstance(tau, torch.Tensor) and tau.dim() > 0:
        om = 1.0 / tau
    else:
        om = torch.full(tuple(f.shape[1:]), 1.0, dtype=f.dtype) / tau
    for k in MRT_STRESS:
        s[k] = om
    mstar = m - s * (m - meq)
    return mrt_inverse(mstar)


def smagorinsky_tau(f: torch.Tensor, tau, cs: float = 0.1, faithful32: bool = True) -> torch.Tensor:
    """Effective tau from the Frobenius norm of the non-equilibrium stress.

    collision.py:187-270.  Note the reference's sqrt(|Pi|^2) has an undefined
    derivative at |Pi| = 0; torch's sqrt backward gives inf*0 = nan there, the
    same hazard as JAX.
    """
    rho = compute_density(f)
    u = compute_velocity(f, rho)
    fneq = f - equilibrium(rho, u, faithful32)
    e = velocities(f.dtype)

    def proj(a, b):
        return torch.einsum("i...,i->...", fneq, e[:, a] * e[:, b])

    sxx, syy, szz = proj(0, 0), proj(1, 1), proj(2, 2)
    sxy, sxz, syz = proj(0, 1), proj(0, 2), proj(1, 2)
    pi = torch.sqrt(sxx * sxx + syy * syy + szz * szz + 2.0 * (sxy * sxy + sxz * sxz + syz * syz))
    return 0.5 * (tau + torch.sqrt(tau * tau + 18.0 * cs * cs * pi / (rho + 1e-10)))


# ---------------------------------------------------------------------------
# Streaming with Bouzidi interpolated bounce-back (jax_lbm/core/streaming.py)
# ---------------------------------------------------------------------------

def stream(f: torch.Tensor, q: torch.Tensor, periodic_yz: bool = True) -> torch.Tensor:
    """Pull streaming plus Bouzidi links.  streaming.py:40-185.

    For direction i at cell x with iopp = opposite(i):
      normal   S_i(x) = P_i(x - e_i), periodic in y/z (or clamped), and at the x
               faces the cell keeps its own P_i(x) (streaming.py:122-131);
      wall     when q[iopp](x) >= 0 (streaming.py:152-153):
               q >= 1/2: P_iopp(x)/(2q) + (1 - 1/(2q)) P_i(x)        (:160-161)
               q <  1/2: 2q P_iopp(x) + (1 - 2q) P_iopp(x + e_i), with the
                         x + e_i neighbour wrapping in all three axes (:166-175)
    """
    nx, ny, nz = f.shape[1], f.shape[2], f.shape[3]
    out = []
    for i in range(Q):
        ex, ey, ez = VELOCITIES[i]
        if i == 0:
            out.append(f[0])
            continue
        fn = f[i]
        if ex:
            fn = torch.roll(fn, ex, 0)
        if ey:
            fn = torch.roll(fn, ey, 1)
        if ez:
            fn = torch.roll(fn, ez, 2)
        # non-periodic x: keep own value where the source falls outside
        xs = torch.arange(nx).reshape(nx, 1, 1)
        if ex == 1:
            fn = torch.where(xs == 0, f[i], fn)
        elif ex == -1:
            fn = torch.where(xs == nx - 1, f[i], fn)
        if not periodic_yz:
            ys = torch.arange(ny).reshape(1, ny, 1)
            zs = torch.arange(nz).reshape(1, 1, nz)
            if ey == 1:
                fn = torch.where(ys == 0, f[i], fn)
            elif ey == -1:
                fn = torch.where(ys == ny - 1, f[i], fn)
            if ez == 1:
                fn = torch.where(zs == 0, f[i], fn)
            elif ez == -1:
                fn = torch.where(zs == nz - 1, f[i], fn)
        io = OPPOSITE[i]
        qo = q[io]
        wall = qo >= 0.0
        inv2q = 1.0 / (2.0 * torch.clamp(qo, 0.5, 1.0))
        hi = inv2q * f[io] + (1.0 - inv2q) * f[i]
        ff = f[io]
        if ex:
            ff = torch.roll(ff, -ex, 0)
        if ey:
            ff = torch.roll(ff, -ey, 1)
        if ez:
            ff = torch.roll(ff, -ez, 2)
        twoq = 2.0 * torch.clamp(qo, 0.0, 0.5)
        lo = twoq * f[io] + (1.0 - twoq) * ff
        bz = torch.where(qo >= 0.5, hi, lo)
        out.append(torch.where(wall, bz, fn))
    return torch.stack(out)


# ---------------------------------------------------------------------------
# Zou-He inlet / outlet (jax_lbm/core/boundary.py)
# ---------------------------------------------------------------------------

def zou_he_inlet(f: torch.Tensor, velocity, rho_target=None) -> torch.Tensor:
    """Velocity inlet at x = 0 over the whole plane.  boundary.py:35-124."""
    ux, uy, uz = velocity[0], velocity[1], velocity[2]
    g = f[:, 0]
    k0 = g[0] + g[3] + g[4] + g[5] + g[6] + g[15] + g[16] + g[17] + g[18]
    km = g[2] + g[8] + g[10] + g[12] + g[14]
    rho_in = rho_target if rho_target is not None else (k0 + 2.0 * km) / (1.0 - ux)
    new = list(g.unbind(0))
    new[1] = g[2] + (1.0 / 3.0) * rho_in * ux
    new[7] = g[10] + (1.0 / 6.0) * rho_in * (ux + uy)
    new[9] = g[8] + (1.0 / 6.0) * rho_in * (ux - uy)
    new[11] = g[14] + (1.0 / 6.0) * rho_in * (ux + uz)
    new[13] = g[12] + (1.0 / 6.0) * rho_in * (ux - uz)
    plane = torch.stack(new).unsqueeze(1)
    return torch.cat([plane, f[:, 1:]], dim=1)


def zou_he_outlet(f: torch.Tensor, rho_out: float = 1.0) -> torch.Tensor:
    """Pressure outlet at x = nx-1 (u_y = u_z = 0 assumed).  boundary.py:127-200."""
    g = f[:, -1]
    k0 = g[0] + g[3] + g[4] + g[5] + g[6] + g[15] + g[16] + g[17] + g[18]
    kp = g[1] + g[7] + g[9] + g[11] + g[13]
    ux = -1.0 + (k0 + 2.0 * kp) / rho_out
    new = list(g.unbind(0))
    new[2] = g[1] - (1.0 / 3.0) * rho_out * ux
    new[10] = g[7] - (1.0 / 6.0) * rho_out * ux
    new[8] = g[9] - (1.0 / 6.0) * rho_out * ux
    new[14] = g[11] - (1.0 / 6.0) * rho_out * ux
    new[12] = g[13] - (1.0 / 6.0) * rho_out * ux
    plane = torch.stack(new).unsqueeze(1)
    return torch.cat([f[:, :-1], plane], dim=1)


# ---------------------------------------------------------------------------
# Forces (jax_lbm/core/forces.py)
# ---------------------------------------------------------------------------

def momentum_exchange(f_post: torch.Tensor, f_streamed: torch.Tensor, q: torch.Tensor) -> torch.Tensor:
    """F = sum over links (q[i](x) >= 0) of e_i (P_i(x) + S_iopp(x)).  forces.py:37-120."""
    e = velocities(f_post.dtype)
    force = torch.zeros(3, dtype=f_post.dtype)
    for i in range(1, Q):
        tot = torch.where(q[i] >= 0.0, f_post[i] + f_streamed[OPPOSITE[i]],
                          torch.zeros((), dtype=f_post.dtype))
        force = force + tot.sum() * e[i]
    return force


def momentum_exchange_decomposed(f_post, f_streamed, q, faithful32: bool = True):
    """(total, pressure) with the pressure part from f_eq(P).  forces.py:123-196."""
    e = velocities(f_post.dtype)
    rho = compute_density(f_post)
    feq = equilibrium(rho, compute_velocity(f_post, rho), faithful32)
    tot = torch.zeros(3, dtype=f_post.dtype)
    pre = torch.zeros(3, dtype=f_post.dtype)
    zero = torch.zeros((), dtype=f_post.dtype)
    for i in range(1, Q):
        wall = q[i] >= 0.0
        io = OPPOSITE[i]
        tot = tot + torch.where(wall, f_post[i] + f_streamed[io], zero).sum() * e[i]
        pre = pre + torch.where(wall, feq[i] + feq[io], zero).sum() * e[i]
    return tot, pre


def projected_area(solid: torch.Tensor, axis: int = 0) -> torch.Tensor:
    """Number of columns along ``axis`` holding a body cell (value 1).  forces.py:199-234."""
    body = solid == 1
    return body.any(dim=axis).to(torch.float32).sum()


def drag_coefficient(force, inlet_velocity, area, rho: float = 1.0):
    """Cd = |F_x| / (0.5 rho U^2 A + 1e-10).  forces.py:237-276."""
    return torch.abs(force[0]) / (0.5 * rho * inlet_velocity * inlet_velocity * area + 1e-10)


# ---------------------------------------------------------------------------
# Geometry (jax_lbm/geometry/{primitives,bouzidi,smooth}.py)
# ---------------------------------------------------------------------------

def world_coords(nx, ny, nz, dtype=torch.float32):
    """Node (i,j,k) sits at ((i+1/2)/s_x - 4, (j+1/2)/s_y - 2, (k+1/2)/s_z - 2)
    with s = (nx/8, ny/4, nz/4).  primitives.py:26-53, smooth.py:62-71."""
    gx = torch.arange(nx, dtype=dtype)
    gy = torch.arange(ny, dtype=dtype)
    gz = torch.arange(nz, dtype=dtype)
    X, Y, Z = torch.meshgrid(gx, gy, gz, indexing="ij")
    sx, sy, sz = nx / 8.0, ny / 4.0, nz / 4.0
    return (X + 0.5) / sx - 4.0, (Y + 0.5) / sy - 2.0, (Z + 0.5) / sz - 2.0, (sx, sy, sz)


def _boundary_links(solid: torch.Tensor, i: int) -> torch.Tensor:
    """fluid cell whose in-bounds neighbour along e_i is a body cell (value 1)."""
    nx, ny, nz = solid.shape
    ex, ey, ez = VELOCITIES[i]
    gx = torch.arange(nx).reshape(nx, 1, 1) + ex
    gy = torch.arange(ny).reshape(1, ny, 1) + ey
    gz = torch.arange(nz).reshape(1, 1, nz) + ez
    inb = (gx >= 0) & (gx < nx) & (gy >= 0) & (gy < ny) & (gz >= 0) & (gz < nz)
    nb = solid[gx.clamp(0, nx - 1), gy.clamp(0, ny - 1), gz.clamp(0, nz - 1)]
    return (solid == 0) & (nb == 1) & inb


def _ray_q(wx, wy, wz, center, radius, i, scale, disc_floor, denom_eps):
    """First ray-sphere hit along e_i (in world units), t in (0, 1] else 1/2,
    clipped to [0.01, 0.99].  primitives.py:115-138, smooth.py:187-203."""
    ex, ey, ez = VELOCITIES[i]
    dt = wx.dtype
    rdx = torch.tensor(float(ex), dtype=dt) / scale[0]
    rdy = torch.tensor(float(ey), dtype=dt) / scale[1]
    rdz = torch.tensor(float(ez), dtype=dt) / scale[2]
    ox, oy, oz = wx - center[0], wy - center[1], wz - center[2]
    a = rdx * rdx + rdy * rdy + rdz * rdz
    b = 2.0 * (ox * rdx + oy * rdy + oz * rdz)
    c = ox * ox + oy * oy + oz * oz - radius * radius
    disc = b * b - 4.0 * a * c
    sq = torch.sqrt(torch.clamp(disc, min=disc_floor))
    t = (-b - sq) / (2.0 * a + denom_eps)
    valid = (disc >= 0.0) & (t > 0.0) & (t <= 1.0)
    qv = torch.where(valid, t, torch.full_like(t, 0.5))
    return torch.clamp(qv, 0.01, 0.99)


def create_sphere(nx, ny, nz, center, radius, coord_dtype=torch.float32):
    """Hard mask (dist^2 <= r^2) and ray-sphere q.  primitives.py:56-174."""
    wx, wy, wz, scale = world_coords(nx, ny, nz, coord_dtype)
    dx, dy, dz = wx - center[0], wy - center[1], wz - center[2]
    solid = (dx * dx + dy * dy + dz * dz <= radius * radius).to(torch.int32)
    q = -torch.ones((Q, nx, ny, nz), dtype=coord_dtype)
    for i in range(1, Q):
        link = _boundary_links(solid, i)
        qv = _ray_q(wx, wy, wz, center, radius, i, scale, 0.0, 1e-20)
        q[i] = torch.where(link, qv, torch.full_like(qv, -1.0))
    return solid, q


def compute_q_from_solid(solid, nx, ny, nz, default_q: float = 0.5, dtype=torch.float32):
    """q = default_q on every fluid->body in-bounds link.  bouzidi.py:32-106."""
    q = -torch.ones((Q, nx, ny, nz), dtype=dtype)
    for i in range(1, Q):
        q[i] = torch.where(_boundary_links(solid, i), torch.full((), default_q, dtype=dtype),
                           torch.full((), -1.0, dtype=dtype))
    return q


def signed_distance_sphere(nx, ny, nz, center, radius, coord_dtype=torch.float32):
    """r - |x_w - c| (positive inside).  smooth.py:34-78."""
    wx, wy, wz, _ = world_coords(nx, ny, nz, coord_dtype)
    dist = torch.sqrt((wx - center[0]) ** 2 + (wy - center[1]) ** 2 + (wz - center[2]) ** 2)
    return radius - dist


def porosity_from_sdf(sdf, sharpness: float = 20.0):
    """sigmoid(sharpness * sdf).  smooth.py:81-115."""
    return torch.sigmoid(sharpness * sdf)


def soft_bounce_back(f_collided, porosity):
    """(1 - phi) f + phi f[opp].  smooth.py:118-155."""
    p = porosity.unsqueeze(0)
    return (1.0 - p) * f_collided + p * f_collided[list(OPPOSITE)]


def smooth_sphere_geometry(nx, ny, nz, center, radius, sharpness: float = 20.0,
                           coord_dtype=torch.float32, disc_floor: float = 1e-10,
                           denom_eps: float = 1e-20):
    """(porosity, solid, q) with solid = porosity > 1/2 and differentiable q(r).

    smooth.py:158-248.  The reference's FP64 KAT driver uses disc_floor 1e-20
    and denom_eps 1e-30 with float64 coordinates (tests/test_gradients_f64.py:145-182).
    """
    wx, wy, wz, scale = world_coords(nx, ny, nz, coord_dtype)
    dist = torch.sqrt((wx - center[0]) ** 2 + (wy - center[1]) ** 2 + (wz - center[2]) ** 2)
    porosity = torch.sigmoid(sharpness * (radius - dist))
    solid = (porosity > 0.5).to(torch.int32)
    qs = [-torch.ones((nx, ny, nz), dtype=porosity.dtype)]
    for i in range(1, Q):
        link = _boundary_links(solid, i)
        qv = _ray_q(wx, wy, wz, center, radius, i, scale, disc_floor, denom_eps)
        qs.append(torch.where(link, qv.to(porosity.dtype), torch.full_like(porosity, -1.0)))
    return porosity, solid, torch.stack(qs)


# ---------------------------------------------------------------------------
# Step composition (jax_lbm/state.py and the examples' scan bodies)
# ---------------------------------------------------------------------------

class LBMState(NamedTuple):
    f: torch.Tensor
    solid: torch.Tensor
    q: torch.Tensor


class LBMParams(NamedTuple):
    tau: object
    inlet_velocity: object
    collision_mode: int = 0
    use_smagorinsky: int = 0
    smagorinsky_cs: float = 0.1
    periodic_yz: bool = True


def _guard_mask(f):
    rho = compute_density(f)
    u = compute_velocity(f, rho)
    speed = torch.sqrt((u * u).sum(dim=0))
    return (rho < 0.3) | (rho > 2.0) | (speed > 0.4) | torch.isnan(rho)


def hot_step(f, *, tau, velocity, q, solid=None, porosity=None, collision="bgk",
             smagorinsky_cs: Optional[float] = None, periodic_yz=True, guard="none",
             f_eq_ref=None, faithful32=True):
    """One timestep, returning (S_{t+1}, P_t).

    Order (state.py:129-171; examples/10_cd_convergence.py:69-83;
    examples/11_optimisation_v2.py:152-160):
      [pre guard: cells with rho<0.3, rho>2, |u|>0.4 or nan(rho) reset to f_eq_ref]
      [tau_eff = smagorinsky_tau(f)] -> bgk|mrt -> hard bounce (solid>=1) or
      soft bounce (porosity) = P_t -> stream(q) -> Zou-He in -> Zou-He out
      [post guard: unstable cells reset to eq(1, 0)]
    """
    if guard == "pre":
        bad = _guard_mask(f)
        f = torch.where(bad.unsqueeze(0), f_eq_ref, f)
    t = tau
    if smagorinsky_cs is not None:
        t = smagorinsky_tau(f, tau, smagorinsky_cs, faithful32)
    fc = mrt(f, t, faithful32) if collision == "mrt" else bgk(f, t, faithful32)
    if porosity is not None:
        fc = soft_bounce_back(fc, porosity)
    elif solid is not None:
        fc = torch.where((solid >= 1).unsqueeze(0), fc[list(OPPOSITE)], fc)
    f_post = fc
    fs = stream(fc, q, periodic_yz)
    fs = zou_he_inlet(fs, velocity)
    fs = zou_he_outlet(fs, 1.0)
    if guard == "post":
        bad = _guard_mask(fs)
        rho = compute_density(fs)
        u = compute_velocity(fs, rho)
        reset = equilibrium(torch.where(bad, torch.ones_like(rho), rho),
                            torch.where(bad.unsqueeze(0), torch.zeros_like(u), u), faithful32)
        fs = torch.where(bad.unsqueeze(0), reset, fs)
    return fs, f_post


def step(state: LBMState, params: LBMParams, faithful32=True) -> LBMState:
    """state.py:87-171 (post guard, hard bounce)."""
    fs, _ = hot_step(state.f, tau=params.tau, velocity=params.inlet_velocity, q=state.q,
                     solid=state.solid, collision="mrt" if params.collision_mode == 1 else "bgk",
                     smagorinsky_cs=params.smagorinsky_cs if params.use_smagorinsky else None,
                     periodic_yz=params.periodic_yz, guard="post", faithful32=faithful32)
    return LBMState(fs, state.solid, state.q)


def simulate(state: LBMState, params: LBMParams, n_steps: int, faithful32=True) -> LBMState:
    """state.py:174-230 (rematerialisation does not change values)."""
    for _ in range(n_steps):
        state = step(state, params, faithful32)
    return state


def rollout(f0, n_steps, **kw):
    """Run ``hot_step`` n_steps times; returns (S_N, P_{N-1}, forces[T,3] or None).

    The per-step force is momentum_exchange(P_t, S_{t+1}, q), the pairing the
    reference's drivers use (tests/test_sphere_cd.py:83-100, examples/10:96-97).
    """
    want_force = kw.pop("forces", False)
    f = f0
    fp = f0
    hist = []
    for _ in range(n_steps):
        f, fp = hot_step(f, **kw)
        if want_force:
            hist.append(momentum_exchange(fp, f, kw["q"]))
    return f, fp, (torch.stack(hist) if want_force else None)


# ---------------------------------------------------------------------------
# Harness protocols of the reference's drivers and KATs
# ---------------------------------------------------------------------------

def channel_walls(nx, ny, nz):
    solid = torch.zeros((nx, ny, nz), dtype=torch.int32)
    solid[:, 0, :] = 1
    solid[:, ny - 1, :] = 1
    return solid


def channel_flow(nx=60, ny=20, nz=4, tau=0.8, u_in=0.05, n_steps=3000, dtype=torch.float32):
    """Poiseuille channel, examples/01_channel_flow.py:22-73: halfway q=1/2 walls at
    y=0, ny-1, f0 = w, BGK, hard bounce, Zou-He in/out, no guard.  Returns the
    z-mean u_x(y) at x = nx//2 and the final state."""
    solid = channel_walls(nx, ny, nz)
    q = compute_q_from_solid(solid, nx, ny, nz, 0.5, dtype=dtype)
    f = equilibrium(torch.ones((nx, ny, nz), dtype=dtype), torch.zeros((3, nx, ny, nz), dtype=dtype))
    vel = torch.tensor([u_in, 0.0, 0.0], dtype=dtype)
    tau_t = torch.tensor(tau, dtype=dtype)
    with torch.no_grad():
        f, _, _ = rollout(f, n_steps, tau=tau_t, velocity=vel, q=q, solid=solid, collision="bgk")
    rho = compute_density(f)
    u = compute_velocity(f, rho)
    return u[0, nx // 2].mean(dim=-1), f


def kat_inlet_velocity_grad(ux=0.05):
    """d mean(rho[2:-2]) / d U after 5 BGK steps on 8x4x4 in FP64.
    tests/test_gradients_f64.py:95-127; published 0.611238 (paper/paper.tex:502)."""
    nx, ny, nz = 8, 4, 4
    U = torch.tensor(ux, dtype=torch.float64, requires_grad=True)
    f = equilibrium(torch.ones((nx, ny, nz), dtype=torch.float64),
                    torch.zeros((3, nx, ny, nz), dtype=torch.float64))
    q = -torch.ones((Q, nx, ny, nz), dtype=torch.float64)
    vel = torch.stack([U, torch.zeros((), dtype=torch.float64), torch.zeros((), dtype=torch.float64)])
    for _ in range(5):
        fc = bgk(f, torch.tensor(0.8, dtype=torch.float64))
        fs = stream(fc, q, True)
        fs = zou_he_inlet(fs, vel)
        f = zou_he_outlet(fs, 1.0)
    loss = compute_density(f[:, 2:-2]).mean()
    (g,) = torch.autograd.grad(loss, U)
    return float(g), float(loss)


def kat_bouzidi_cd(radius, nx=32, ny=16, nz=16, n_steps=10, u_in=0.05):
    """Cd(r) of tests/test_gradients_f64.py:129-207 in FP64: soft porosity
    (alpha=20), ray q(r), BGK tau=3 nu + 1/2 with nu = U*8/50, init eq(1, U),
    Zou-He, Cd from the last (P, S) pair, area = max(A, 1)."""
    dt = torch.float64
    D = 32 * 0.5 / 4.0
    tau = torch.tensor(3.0 * (u_in * D / 50.0) + 0.5, dtype=dt)
    porosity, solid, q = smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), radius, 20.0,
                                                coord_dtype=dt, disc_floor=1e-20, denom_eps=1e-30)
    f = equilibrium(torch.ones((nx, ny, nz), dtype=dt),
                    torch.stack([torch.full((nx, ny, nz), u_in, dtype=dt),
                                 torch.zeros((nx, ny, nz), dtype=dt),
                                 torch.zeros((nx, ny, nz), dtype=dt)]))
    vel = torch.tensor([u_in, 0.0, 0.0], dtype=dt)
    fp = f
    for _ in range(n_steps):
        fc = bgk(f, tau)
        fc = soft_bounce_back(fc, porosity)
        fp = fc
        fs = stream(fc, q, True)
        fs = zou_he_inlet(fs, vel)
        f = zou_he_outlet(fs, 1.0)
    force = momentum_exchange(fp, f, q)
    area = torch.clamp(projected_area(solid, 0).to(dt), min=1.0)
    return drag_coefficient(force, u_in, area)


def kat_bouzidi_grad(r0=0.5):
    r = torch.tensor(r0, dtype=torch.float64, requires_grad=True)
    cd = kat_bouzidi_cd(r)
    (g,) = torch.autograd.grad(cd, r)
    return float(g), float(cd)


def sphere_center_world(nx, ny, nz, center_lattice):
    """Continuous lattice coordinate X (node i at X = i + 1/2) -> world."""
    sx, sy, sz = nx / 8.0, ny / 4.0, nz / 4.0
    return (center_lattice[0] / sx - 4.0, center_lattice[1] / sy - 2.0, center_lattice[2] / sz - 2.0)


def adam_update(r, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """optax.adam (scale_by_adam + scale(-lr)) for one scalar, t counted from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return r - lr * mh / (math.sqrt(vh) + eps), m, v
