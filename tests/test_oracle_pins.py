"""Pins of the oracle's collision closures and force decomposition that the
published known answers leave open (the published KATs are all BGK or
pure momentum exchange).

* smagorinsky_tau (collision.py:187-270): Chapman-Enskog closed form.  For
  f = f_eq(rho, u) + f1 with f1_i = -tau_t w_i rho / c_s^2 Q_i:S (Q_i = e_i e_i
  - c_s^2 I, S a symmetric strain), the non-equilibrium stress is
  Pi = -2 tau_t rho c_s^2 S exactly on D3Q19 (fourth-order isotropy), and the
  fixed point of tau_eff = (tau + sqrt(tau^2 + 18 C^2 |Pi|_F / rho)) / 2 with
  tau_t = tau_eff is tau_eff = tau + 3 C^2 |S|_F.  A hand-built pure shear
  pins the Frobenius norm's off-diagonal factor.
* mrt (collision.py:103-184, lattice.py:106-129): the shear viscosity of a
  decaying shear wave is (tau - 1/2)/3 in every orientation, which pins which
  moment slots carry omega = 1/tau; with every rate set to omega, MRT equals
  BGK (pins M, M^-1 and m_eq together).
* momentum_exchange_decomposed (forces.py:123-196): the reference's
  invariant (tests/test_forces.py:45-56, friction = 0 at equilibrium) and a
  closed form of the pressure part for a plane of links.

Each pin is shown to reject a plausible mistake planted in the oracle.
"""

import math

import torch

from oracle import lbm_oracle as O

F64 = torch.float64


# ------------------------------------------------------------ Smagorinsky --

def _strain_state(rho, u, S, tau_t):
    """f_eq(rho, u) + the first-order Chapman-Enskog term for strain S."""
    e = O.velocities(F64)
    w = O.weights(F64, faithful32=False)
    n = rho.shape
    Qs = torch.einsum("ia,ib,ab...->i...", e, e, S) - (S.diagonal(0, 0, 1).sum(-1) / 3.0).unsqueeze(0)
    f1 = -tau_t * w.reshape((19,) + (1,) * len(n)) * rho.unsqueeze(0) * 3.0 * Qs
    return O.equilibrium(rho, u, faithful32=False) + f1


def _smag_closed_form_error(fn):
    g = torch.Generator().manual_seed(11)
    shape = (3, 4, 2)
    rho = 1.0 + 0.05 * torch.randn(shape, generator=g, dtype=F64)
    u = 0.04 * torch.randn((3,) + shape, generator=g, dtype=F64)
    A = 2e-3 * torch.randn((3, 3) + shape, generator=g, dtype=F64)
    S = 0.5 * (A + A.transpose(0, 1))  # symmetric, with a trace
    tau, cs = 0.56, 0.1
    snorm = torch.sqrt((S * S).sum((0, 1)))
    expect = tau + 3.0 * cs * cs * snorm
    f = _strain_state(rho, u, S, expect)
    got = fn(f, tau, cs, faithful32=False)
    # relative effect of the reference's +1e-10 on rho is far below the bound
    return float(((got - expect) / (expect - tau)).abs().max())


def test_smagorinsky_chapman_enskog_closed_form():
    assert _smag_closed_form_error(O.smagorinsky_tau) < 1e-7


def _smag_planted(f, tau, cs=0.1, faithful32=True):
    """Plant: 9 C^2 instead of 18 C^2 and no factor 2 on the off-diagonal stresses."""
    rho = O.compute_density(f)
    u = O.compute_velocity(f, rho)
    fneq = f - O.equilibrium(rho, u, faithful32)
    e = O.velocities(f.dtype)
    p = lambda a, b: torch.einsum("i...,i->...", fneq, e[:, a] * e[:, b])  # noqa: E731
    pi = torch.sqrt(p(0, 0) ** 2 + p(1, 1) ** 2 + p(2, 2) ** 2 + p(0, 1) ** 2 + p(0, 2) ** 2 + p(1, 2) ** 2)
    return 0.5 * (tau + torch.sqrt(tau * tau + 9.0 * cs * cs * pi / (rho + 1e-10)))


def test_smagorinsky_planted_mistake_rejected():
    assert _smag_closed_form_error(_smag_planted) > 1e-2


def test_smagorinsky_pure_shear_by_hand():
    """f7, f10 += d and f8, f9 -= d leave rho and u unchanged and give
    Pi_xy = Pi_yx = 4 d (all else 0), so |Pi|_F = 4 sqrt(2) d."""
    d, tau, cs = 1e-3, 0.7, 0.1
    f = O.equilibrium(torch.ones(1, dtype=F64), torch.zeros((3, 1), dtype=F64), faithful32=False)
    f[7] += d
    f[10] += d
    f[8] -= d
    f[9] -= d
    got = float(O.smagorinsky_tau(f, tau, cs, faithful32=False)[0])
    pi = 4.0 * math.sqrt(2.0) * d
    expect = 0.5 * (tau + math.sqrt(tau * tau + 18.0 * cs * cs * pi / (1.0 + 1e-10)))
    assert abs(got - expect) < 1e-14
    bad = float(_smag_planted(f, tau, cs, faithful32=False)[0])
    assert abs(bad - expect) > 1e-4


# ------------------------------------------------------------------- MRT --

def _stream_periodic(f):
    """Pull streaming with every axis periodic (textbook; no walls)."""
    out = torch.empty_like(f)
    for i, (ex, ey, ez) in enumerate(O.VELOCITIES):
        out[i] = torch.roll(f[i], shifts=(ex, ey, ez), dims=(0, 1, 2))
    return out


ORIENTATIONS = {
    # name: (grid, wave vector direction, velocity direction); stress slots excited
    "k_y_u_x": ((1, 64, 1), (0, 1, 0), (1, 0, 0)),     # p_xy (13)
    "k_z_u_y": ((1, 1, 64), (0, 0, 1), (0, 1, 0)),     # p_yz (14)
    "k_z_u_x": ((1, 1, 64), (0, 0, 1), (1, 0, 0)),     # p_xz (15)
    "k_yz_u_y-z": ((1, 64, 64), (0, 1, 1), (0, 1, -1)),  # p_ww (11)
    "k_xy_u_x-y": ((64, 64, 1), (1, 1, 0), (1, -1, 0)),  # 3p_xx (9) and p_ww (11)
    "k_xz_u_x-z": ((64, 1, 64), (1, 0, 1), (1, 0, -1)),  # 3p_xx (9) and p_ww (11)
}


def _shear_wave_viscosity(collide, name, tau=0.8, amp=1e-3, t0=20, t1=120):
    (nx, ny, nz), kd, vd = ORIENTATIONS[name]
    L = max(nx, ny, nz)
    kv = torch.tensor(kd, dtype=F64) * (2.0 * math.pi / L)
    v = torch.tensor(vd, dtype=F64)
    v = v / v.norm()
    X, Y, Z = torch.meshgrid(torch.arange(nx, dtype=F64), torch.arange(ny, dtype=F64),
                             torch.arange(nz, dtype=F64), indexing="ij")
    phase = kv[0] * X + kv[1] * Y + kv[2] * Z
    mode = torch.sin(phase)
    u = amp * v.reshape(3, 1, 1, 1) * mode
    f = O.equilibrium(torch.ones((nx, ny, nz), dtype=F64), u, faithful32=False)
    e = O.velocities(F64)

    def amplitude(f):
        j = torch.einsum("ixyz,ia->axyz", f, e)
        return float((j * v.reshape(3, 1, 1, 1) * mode).sum() / (mode * mode).sum())

    a0 = None
    for t in range(t1 + 1):
        if t == t0:
            a0 = amplitude(f)
        if t == t1:
            a1 = amplitude(f)
            break
        f = _stream_periodic(collide(f, tau))
    if a0 <= 0.0 or a1 <= 0.0:
        return math.inf  # the wave changed sign: not a viscous decay at all
    return math.log(a0 / a1) / ((t1 - t0) * float(kv @ kv))


def test_mrt_shear_viscosity_isotropic():
    """nu = (tau - 1/2)/3 for shear waves along the axes and the diagonals,
    to the O(k^2) accuracy of the scheme: 9e-4 (axes) and 5e-4 (diagonals)
    at 64 cells per wavelength, 4x that at 32."""
    tau = 0.8
    nu = (tau - 0.5) / 3.0
    for name in ORIENTATIONS:
        got = _shear_wave_viscosity(lambda f, t: O.mrt(f, t, faithful32=False), name, tau)
        assert abs(got / nu - 1.0) < 2e-3, (name, got, nu)


def test_mrt_shear_viscosity_catches_wrong_stress_slots(monkeypatch):
    """Plant: omega in slot 10 instead of 9 (and 12 instead of 11)."""
    tau = 0.8
    nu = (tau - 0.5) / 3.0
    for wrong in ((10, 11, 13, 14, 15), (9, 12, 13, 14, 15)):
        monkeypatch.setattr(O, "MRT_STRESS", wrong)
        errs = [abs(_shear_wave_viscosity(lambda f, t: O.mrt(f, t, faithful32=False), n, tau) / nu - 1.0)
                for n in ("k_yz_u_y-z", "k_xy_u_x-y")]
        assert max(errs) > 5e-2, (wrong, errs)
    monkeypatch.undo()


def test_bgk_shear_viscosity():
    tau = 0.8
    got = _shear_wave_viscosity(lambda f, t: O.bgk(f, t, faithful32=False), "k_xy_u_x-y", tau)
    assert abs(got / ((tau - 0.5) / 3.0) - 1.0) < 2e-3


def test_mrt_with_every_rate_omega_is_bgk(monkeypatch):
    g = torch.Generator().manual_seed(21)
    rho = 1.0 + 0.02 * torch.randn((4, 3, 2), generator=g, dtype=F64)
    u = 0.05 * torch.randn((3, 4, 3, 2), generator=g, dtype=F64)
    f = O.equilibrium(rho, u, faithful32=False) + 1e-3 * torch.randn((19, 4, 3, 2), generator=g, dtype=F64)
    tau = 0.63
    monkeypatch.setattr(O, "mrt_rates_fixed", lambda dtype=F64, faithful32=True: torch.full((19,), 1.0 / tau,
                                                                                           dtype=dtype))
    got = O.mrt(f, tau, faithful32=False)
    assert float((got - O.bgk(f, tau, faithful32=False)).abs().max()) < 1e-14
    # the same check rejects a wrong row norm in the inverse transform
    # (plant: ||M_10||^2 = 36 instead of 72)
    monkeypatch.setattr(O, "MRT_ROW_NORMS_REF", O.MRT_ROW_NORMS_REF[:10] + (36.0,) + O.MRT_ROW_NORMS_REF[11:])
    bad = O.mrt(f, tau, faithful32=False)
    assert float((bad - O.bgk(f, tau, faithful32=False)).abs().max()) > 1e-8


# ------------------------------------------------- force decomposition --

def _plane_links(nx=16, ny=8, nz=8):
    q = -torch.ones((19, nx, ny, nz), dtype=F64)
    q[1, 7] = 0.5  # direction 1 from the x = 7 plane hits a wall (test_forces.py:41)
    return q


def test_decomposition_friction_zero_at_equilibrium():
    """tests/test_forces.py:45-56: total - pressure = 0 when f_post = f_streamed = f_eq."""
    nx, ny, nz = 16, 8, 8
    u = torch.zeros((3, nx, ny, nz), dtype=F64)
    u[0] = 0.05
    f = O.equilibrium(torch.ones((nx, ny, nz), dtype=F64), u, faithful32=False)
    total, pressure = O.momentum_exchange_decomposed(f, f, _plane_links(), faithful32=False)
    # the reference asserts 5e-4 in float32; in float64 with exact weights it is rounding
    assert float((total - pressure).abs().max()) < 1e-12
    assert float(total[0]) > 1.0


def _pressure_closed_form_error(fn):
    """P = f_eq(rho0, (U,0,0)) + a zero-moment shear perturbation, S random:
    the pressure part on the plane of 64 direction-1 links is
    64 (f_eq_1 + f_eq_2)(P) = 64 * 2 w_1 rho0 (1 + 3 U^2) along x, whatever S."""
    nx, ny, nz = 16, 8, 8
    rho0, U = 1.02, 0.05
    u = torch.zeros((3, nx, ny, nz), dtype=F64)
    u[0] = U
    P = O.equilibrium(torch.full((nx, ny, nz), rho0, dtype=F64), u, faithful32=False)
    P[7] += 1e-3
    P[10] += 1e-3
    P[8] -= 1e-3
    P[9] -= 1e-3
    g = torch.Generator().manual_seed(4)
    S = P + 1e-2 * torch.rand(P.shape, generator=g, dtype=F64)
    q = _plane_links()
    total, pressure = fn(P, S, q)
    expect_p = 64.0 * 2.0 / 18.0 * rho0 * (1.0 + 3.0 * U * U)
    expect_t = float((P[1, 7] + S[2, 7]).sum())
    return max(abs(float(pressure[0]) - expect_p), float(pressure[1:].abs().max()), abs(float(total[0]) - expect_t))


def test_decomposition_pressure_closed_form():
    assert _pressure_closed_form_error(lambda P, S, q: O.momentum_exchange_decomposed(P, S, q, faithful32=False)) \
        < 1e-12


def test_decomposition_planted_mistakes_rejected():
    def eq_of_streamed(P, S, q):  # plant: the splitting equilibrium taken from f_streamed
        return O.momentum_exchange_decomposed(S, S, q, faithful32=False)[1], None

    def same_direction_twice(P, S, q):  # plant: f_eq_i + f_eq_i instead of f_eq_i + f_eq_iopp
        e = O.velocities(F64)
        rho = O.compute_density(P)
        feq = O.equilibrium(rho, O.compute_velocity(P, rho), faithful32=False)
        tot, pre = torch.zeros(3, dtype=F64), torch.zeros(3, dtype=F64)
        for i in range(1, 19):
            wall = q[i] >= 0
            tot = tot + torch.where(wall, P[i] + S[O.OPPOSITE[i]], torch.zeros((), dtype=F64)).sum() * e[i]
            pre = pre + torch.where(wall, 2.0 * feq[i], torch.zeros((), dtype=F64)).sum() * e[i]
        return tot, pre

    def wrap(fn):
        def g(P, S, q):
            t, p = fn(P, S, q)
            if p is None:
                return O.momentum_exchange(P, S, q), t
            return t, p
        return g

    assert _pressure_closed_form_error(wrap(eq_of_streamed)) > 1e-4
    assert _pressure_closed_form_error(same_direction_twice) > 1e-4
