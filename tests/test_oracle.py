"""Pins of the CPU oracle: published known answers, closed forms and
planted-mistake checks (each plants an error in the oracle and shows a pin
catches it)."""

import json
import math
import os

import pytest
import torch

from oracle import lbm_oracle as O

# values the reference publishes for this path, with their sources
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "kats.json")) as _fh:
    KAT = json.load(_fh)


# ---------------------------------------------------------------- lattice --

def test_weight_moments_closed_form():
    w = O.weights(torch.float64, faithful32=False)
    e = O.velocities(torch.float64)
    assert abs(float(w.sum()) - 1.0) < 1e-15
    assert float((w[:, None] * e).sum(0).abs().max()) < 1e-15
    second = torch.einsum("i,ia,ib->ab", w, e, e)
    assert torch.allclose(second, torch.eye(3, dtype=torch.float64) / 3.0, atol=1e-15)
    fourth_xxyy = float((w * e[:, 0] ** 2 * e[:, 1] ** 2).sum())
    assert abs(fourth_xxyy - 1.0 / 9.0) < 1e-15  # isotropy: c_s^4


def test_opposite_table():
    assert O.OPPOSITE == (0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15)


def test_mrt_matrix_orthogonal_with_reference_norms():
    M = O.mrt_matrix(torch.float64)
    G = M @ M.T
    assert torch.allclose(G, torch.diag(torch.tensor(O.MRT_ROW_NORMS_REF, dtype=torch.float64)))


def test_mrt_roundtrip_and_equilibrium_moments():
    g = torch.Generator().manual_seed(42)
    f = torch.rand(19, generator=g, dtype=torch.float64)
    assert torch.allclose(O.mrt_inverse(O.mrt_forward(f)), f, atol=1e-14)
    rho = torch.tensor(1.03, dtype=torch.float64)
    u = torch.tensor([0.05, -0.02, 0.01], dtype=torch.float64)
    feq = O.equilibrium(rho.reshape(1), u.reshape(3, 1), faithful32=False)[:, 0]
    # m_eq(rho, u) = M f_eq(rho, u) (test_collision.py:57-70)
    assert torch.allclose(O.mrt_forward(feq), O.mrt_equilibrium_moments(rho, u), atol=1e-14)


# ---------------------------------------------------------- equilibrium --

def test_equilibrium_moments_closed_form():
    g = torch.Generator().manual_seed(0)
    rho = 1 + 0.1 * torch.rand((3, 2, 2), generator=g, dtype=torch.float64)
    u = 0.05 * torch.randn((3, 3, 2, 2), generator=g, dtype=torch.float64)
    f = O.equilibrium(rho, u, faithful32=False)
    e = O.velocities(torch.float64)
    assert torch.allclose(f.sum(0), rho, atol=1e-14)
    assert torch.allclose(torch.einsum("ixyz,ia->axyz", f, e), rho * u, atol=1e-14)
    pi = torch.einsum("ixyz,ia,ib->abxyz", f, e, e)
    expect = rho * (torch.eye(3, dtype=torch.float64)[:, :, None, None, None] / 3.0
                    + torch.einsum("axyz,bxyz->abxyz", u, u))
    assert torch.allclose(pi, expect, atol=1e-14)


def test_kat_equilibrium_mass_gradient():
    """d(sum f_eq)/d rho = N_cells = 64 on 4^3 (paper/paper.tex:501)."""
    rho = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    f = O.equilibrium(rho.expand(4, 4, 4), torch.zeros((3, 4, 4, 4), dtype=torch.float64))
    (g,) = torch.autograd.grad(f.sum(), rho)
    assert abs(float(g) - KAT["d_sum_feq_d_rho"]["value"]) < 1e-5


# ------------------------------------------------------------ collision --

@pytest.mark.parametrize("op", ["bgk", "mrt"])
def test_collision_fixed_point_and_conservation(op):
    rho = torch.ones((4, 4, 4), dtype=torch.float64)
    u = torch.zeros((3, 4, 4, 4), dtype=torch.float64)
    u[0] = 0.05
    feq = O.equilibrium(rho, u)
    assert torch.allclose(getattr(O, op)(feq, 0.8), feq, atol=1e-14)
    g = torch.Generator().manual_seed(1)
    f = feq + 1e-3 * torch.randn(feq.shape, generator=g, dtype=torch.float64)
    out = getattr(O, op)(f, 0.8)
    e = O.velocities(torch.float64)
    assert torch.allclose(out.sum(0), f.sum(0), atol=1e-14)
    assert torch.allclose(torch.einsum("ixyz,ia->axyz", out, e), torch.einsum("ixyz,ia->axyz", f, e), atol=1e-14)


def test_smagorinsky_bounds():
    rho = torch.ones((4, 4, 4), dtype=torch.float64)
    u = torch.zeros((3, 4, 4, 4), dtype=torch.float64)
    u[0] = 0.05
    feq = O.equilibrium(rho, u)
    assert torch.allclose(O.smagorinsky_tau(feq, 0.8), torch.full((4, 4, 4), 0.8, dtype=torch.float64), atol=1e-7)
    g = torch.Generator().manual_seed(3)
    f = feq + 1e-3 * torch.randn(feq.shape, generator=g, dtype=torch.float64)
    assert bool((O.smagorinsky_tau(f, 0.6) >= 0.6 - 1e-12).all())


# ------------------------------------------------------------ streaming --

def test_stream_shift_and_rest():
    f = torch.zeros((19, 8, 8, 8), dtype=torch.float64)
    f[3, :, 3, :] = 1.0
    f[0] = torch.arange(512, dtype=torch.float64).reshape(8, 8, 8)
    out = O.stream(f, -torch.ones_like(f))
    assert torch.equal(out[3, :, 4, :], torch.ones((8, 8), dtype=torch.float64))
    assert torch.equal(out[3, :, 3, :], torch.zeros((8, 8), dtype=torch.float64))
    assert torch.equal(out[0], f[0])


def test_stream_bouzidi_branches_continuous_at_half():
    g = torch.Generator().manual_seed(5)
    f = torch.rand((19, 6, 5, 4), generator=g, dtype=torch.float64)
    for qv in (0.5, 0.5 - 1e-12):
        q = -torch.ones_like(f)
        q[2, 3] = qv
        out = O.stream(f, q)
        assert abs(float(out[1, 3, 2, 1]) - float(f[2, 3, 2, 1])) < 1e-10


def test_stream_mass_conserved_periodic():
    """No walls, x-invariant field: each population's total is conserved
    exactly (streaming is a permutation; the x-face fallback equals the value
    that would have arrived).  test_streaming.py:20-31."""
    g = torch.Generator().manual_seed(6)
    f = torch.rand((19, 1, 5, 4), generator=g, dtype=torch.float64).expand(19, 6, 5, 4).contiguous()
    out = O.stream(f, -torch.ones_like(f))
    assert torch.allclose(out.sum((1, 2, 3)), f.sum((1, 2, 3)), atol=1e-12)
    # y/z shifts only: an x-invariant field pulled along +y moves by one row
    assert torch.allclose(out[3, :, 1], f[3, :, 0])


# ------------------------------------------------------------- Zou-He --

def test_zou_he_post_conditions():
    g = torch.Generator().manual_seed(7)
    f = O.equilibrium(torch.ones((5, 3, 3), dtype=torch.float64), torch.zeros((3, 5, 3, 3), dtype=torch.float64))
    f = f + 1e-3 * torch.rand(f.shape, generator=g, dtype=torch.float64)
    vel = torch.tensor([0.05, 0.0, 0.0], dtype=torch.float64)
    out = O.zou_he_inlet(f, vel)
    e = O.velocities(torch.float64)
    rho_in = out[:, 0].sum(0)
    assert torch.allclose((out[:, 0] * e[:, 0].reshape(19, 1, 1)).sum(0), rho_in * 0.05, atol=1e-15)
    out = O.zou_he_outlet(f, 1.0)
    assert torch.allclose(out[:, -1].sum(0), torch.ones((3, 3), dtype=torch.float64), atol=1e-15)


# -------------------------------------------------------------- forces --

def test_forces_known_values():
    k = KAT["drag_coefficient"]
    cd = O.drag_coefficient(torch.tensor([k["force_x"], 0.0, 0.0]), k["u"], k["area"])
    assert abs(float(cd) - k["value"]) < 1e-6
    k = KAT["projected_area_box"]
    solid = torch.zeros(k["shape"], dtype=torch.int32)
    (x0, x1), (y0, y1), (z0, z1) = k["box"]
    solid[x0:x1, y0:y1, z0:z1] = 1
    assert float(O.projected_area(solid, 0)) == k["value"]
    ground = torch.zeros((16, 8, 8), dtype=torch.int32)
    ground[:, 0, :] = 2
    assert float(O.projected_area(ground, 0)) == KAT["projected_area_ground"]["value"]
    f = torch.rand((19, 4, 4, 4), dtype=torch.float64)
    assert float(O.momentum_exchange(f, f, -torch.ones_like(f)).abs().max()) == 0.0


# ------------------------------------------------------------- geometry --

def test_ray_q_derivative_closed_form():
    """dq/dr = -2 r / sqrt(disc) (world units) where q is unclipped and valid."""
    nx, ny, nz = 32, 16, 16
    r = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    _, solid, q = O.smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), r, 20.0, coord_dtype=torch.float64)
    i = 1
    mask = (q[i] > 0.01) & (q[i] < 0.99)
    idx = mask.nonzero()[0]
    (g,) = torch.autograd.grad(q[i][tuple(idx)], r)
    wx, wy, wz, scale = O.world_coords(nx, ny, nz, torch.float64)
    ox, oy, oz = (float(t[tuple(idx)]) for t in (wx, wy, wz))
    a = (1.0 / scale[0]) ** 2
    b = 2.0 * ox / scale[0]
    c = ox * ox + oy * oy + oz * oz - 0.25
    disc = b * b - 4 * a * c
    assert abs(float(g) - (-2.0 * 0.5 / math.sqrt(disc))) < 1e-9


def test_smooth_solid_matches_create_sphere_without_ties():
    nx, ny, nz = 32, 16, 16
    s_hard, _ = O.create_sphere(nx, ny, nz, (0.0, 0.0, 0.0), 0.5)
    _, s_soft, _ = O.smooth_sphere_geometry(nx, ny, nz, (0.0, 0.0, 0.0), 0.5, 20.0)
    assert torch.equal(s_hard, s_soft)


# ------------------------------------------------------- published KATs --

def test_kat_inlet_velocity_gradient():
    """FP64 d mean(rho)/dU over 5 steps = 0.611238 (paper/paper.tex:502)."""
    g, _ = O.kat_inlet_velocity_grad()
    assert abs(g - KAT["d_mean_rho_d_u_inlet"]["value"]) < 5e-7


def test_kat_bouzidi_radius_gradient():
    """FP64 Bouzidi dCd/dr over 10 steps = 93.33275 (paper/paper.tex:503)."""
    g, _ = O.kat_bouzidi_grad()
    assert abs(g - KAT["d_cd_d_r_bouzidi_soft"]["value"]) < 5e-5


def test_kat_poiseuille_profile():
    """60x20x4 BGK tau=0.8, 3000 steps, FP32: README.md:59-69 to 6 decimals."""
    prof, _ = O.channel_flow()
    for y, v in KAT["poiseuille_60x20x4"]["ux_by_y"].items():
        assert round(float(prof[int(y)]), 6) == v, (y, float(prof[int(y)]))


# ---------------------------------------------------- planted mistakes --

def test_planted_mistakes_are_caught(monkeypatch):
    """A wrong Bouzidi branch or a wrong Zou-He term moves a published KAT."""
    base, _ = O.kat_bouzidi_grad()
    orig_stream = O.stream

    def wrong_stream(f, q, periodic_yz=True):
        # plant: use q instead of 2q in the near-wall branch
        return orig_stream(f, torch.where((q >= 0) & (q < 0.5), 0.5 * q, q), periodic_yz)

    monkeypatch.setattr(O, "stream", wrong_stream)
    bad, _ = O.kat_bouzidi_grad()
    monkeypatch.setattr(O, "stream", orig_stream)
    assert abs(bad - 93.33275) > 1e-2 and abs(base - 93.33275) < 5e-5

    orig_in = O.zou_he_inlet

    def wrong_inlet(f, velocity, rho_target=None):
        # plant: f1 = f2 + rho u_x / 6 instead of / 3 (boundary.py:118)
        out = orig_in(f, velocity, rho_target)
        g = f[:, 0]
        k0 = g[0] + g[3] + g[4] + g[5] + g[6] + g[15] + g[16] + g[17] + g[18]
        km = g[2] + g[8] + g[10] + g[12] + g[14]
        rho_in = (k0 + 2.0 * km) / (1.0 - velocity[0])
        plane = out[:, 0].clone()
        plane[1] = g[2] + rho_in * velocity[0] / 6.0
        return torch.cat([plane.unsqueeze(1), out[:, 1:]], dim=1)

    monkeypatch.setattr(O, "zou_he_inlet", wrong_inlet)
    g_bad, _ = O.kat_inlet_velocity_grad()
    monkeypatch.setattr(O, "zou_he_inlet", orig_in)
    assert abs(g_bad - 0.611238) > 1e-3

    orig_w = O.weights

    def wrong_weights(dtype=torch.float64, faithful32=True):
        w = orig_w(dtype, faithful32).clone()
        w[7], w[8] = w[8] * 1.01, w[7] * 0.99  # plant: perturbed edge weights
        return w

    monkeypatch.setattr(O, "weights", wrong_weights)
    prof, _ = O.channel_flow(n_steps=300)
    monkeypatch.setattr(O, "weights", orig_w)
    prof_ok, _ = O.channel_flow(n_steps=300)
    assert float((prof - prof_ok).abs().max()) > 1e-5
