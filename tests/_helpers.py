"""Shared helpers for the parity tests (oracle on the CPU, kernels on the GPU)."""

import torch

from oracle import lbm_oracle as O


def near_eq_state(nx, ny, nz, dtype=torch.float64, seed=0, amp=1e-3, u=(0.05, 0.01, -0.02)):
    """Equilibrium at rho~1, u~u plus seeded noise, shape (19, nx, ny, nz) on the CPU."""
    g = torch.Generator().manual_seed(seed)
    rho = 1.0 + 0.01 * torch.randn((nx, ny, nz), generator=g, dtype=torch.float64)
    uu = torch.stack([torch.full((nx, ny, nz), v, dtype=torch.float64) for v in u])
    uu = uu + 0.01 * torch.randn((3, nx, ny, nz), generator=g, dtype=torch.float64)
    f = O.equilibrium(rho, uu)
    f = f + amp * torch.randn(f.shape, generator=g, dtype=torch.float64)
    return f.to(dtype)


def random_q(nx, ny, nz, seed=1, p_wall=0.15, dtype=torch.float64):
    """q with about p_wall links per direction, half in [0.01,0.5), half in [0.5,0.99]."""
    g = torch.Generator().manual_seed(seed)
    u = torch.rand((19, nx, ny, nz), generator=g, dtype=torch.float64)
    v = torch.rand((19, nx, ny, nz), generator=g, dtype=torch.float64)
    q = torch.where(u < p_wall, 0.01 + 0.98 * v, torch.full_like(v, -1.0))
    q[0] = -1.0
    return q.to(dtype)


def gpu(t, dev):
    return t.to(dev)


def cpu(t):
    return t.detach().to("cpu")


def maxabs(a, b):
    return float((cpu(a).to(torch.float64) - cpu(b).to(torch.float64)).abs().max())


def perturbed_start(nx, ny, nz, seed, amp=1e-8):
    """f_0 = w_i + amp * N(0, 1), the noise drawn in float64 on the CPU from
    torch.Generator().manual_seed(seed) with shape (19, nx, ny, nz)."""
    g = torch.Generator().manual_seed(int(seed))
    w = O.weights(torch.float64).reshape(19, 1, 1, 1)
    return w + amp * torch.randn((19, nx, ny, nz), generator=g, dtype=torch.float64)
