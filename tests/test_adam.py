"""The radius optimiser of the sphere-drag study, checked on the CPU.

The reference runs optax.adam(lr) on the radius and clips it to
[r_min, r_max] after each update (examples/11_optimisation_v2.py:182-194).
optax is not installed, so the update is pinned by what fixes the published
algorithm (Kingma & Ba, "Adam", Algorithm 1, with eps added outside the
square root and no eps inside it, as optax.scale_by_adam does by default):

1. torch.optim.Adam implements the same update; with the clip bounds out of
   reach, both produce the same radius sequence for any gradient sequence.
2. With a constant gradient g the bias-corrected moments are exactly g and
   g^2 at every iteration, so each update is -lr * g / (|g| + eps); a missing
   bias correction moves the first step by a factor of about 3.16.
3. The clip holds r inside [r_min, r_max] and the history records the radius
   each gradient was evaluated at.
"""

import math

import pytest
import torch

from lbm_jax_autodiff_amd.protocols import adam


def _quartic(r):
    # a smooth non-quadratic loss, so the gradient sequence varies in size and sign
    loss = (r - 0.41) ** 4 + 0.3 * math.sin(7.0 * r)
    return loss, 4.0 * (r - 0.41) ** 3 + 2.1 * math.cos(7.0 * r)


@pytest.mark.parametrize("lr", [0.05, 0.01])
def test_adam_matches_torch_adam(lr):
    n = 30
    r_ours, hist = adam(_quartic, 0.5, n, lr, r_min=-10.0, r_max=10.0)
    p = torch.nn.Parameter(torch.tensor([0.5], dtype=torch.float64))
    opt = torch.optim.Adam([p], lr=lr, betas=(0.9, 0.999), eps=1e-8)
    for t in range(n):
        assert hist[t]["r"] == pytest.approx(float(p), rel=0, abs=1e-12)
        opt.zero_grad()
        p.grad = torch.tensor([_quartic(float(p))[1]], dtype=torch.float64)
        opt.step()
    assert r_ours == pytest.approx(float(p), rel=0, abs=1e-12)
    # the run moved the radius by several steps and met gradients of both signs
    rs = [h["r"] for h in hist]
    assert max(rs) - min(rs) > 3 * lr
    assert min(h["grad"] for h in hist) < 0.0 < max(h["grad"] for h in hist)


def test_oracle_adam_update_matches_torch_adam():
    from oracle import lbm_oracle as O
    r, m, v = 0.5, 0.0, 0.0
    p = torch.nn.Parameter(torch.tensor([0.5], dtype=torch.float64))
    opt = torch.optim.Adam([p], lr=0.05, betas=(0.9, 0.999), eps=1e-8)
    for t in range(1, 26):
        g = _quartic(r)[1]
        r, m, v = O.adam_update(r, g, m, v, t, 0.05)
        p.grad = torch.tensor([g], dtype=torch.float64)
        opt.step()
        assert r == pytest.approx(float(p), rel=0, abs=1e-12)


@pytest.mark.parametrize("g", [3.4, -0.02, 1e-7])
def test_adam_constant_gradient_closed_form(g):
    lr, eps = 0.05, 1e-8
    r_end, hist = adam(lambda r: (0.0, g), 0.5, 6, lr, r_min=-10.0, r_max=10.0, eps=eps)
    step = lr * g / (abs(g) + eps)
    for t, h in enumerate(hist):
        assert h["r"] == pytest.approx(0.5 - t * step, rel=0, abs=1e-13)
    assert r_end == pytest.approx(0.5 - 6 * step, rel=0, abs=1e-13)


def test_adam_clip_and_history():
    # a positive gradient drives r down by about lr per iteration onto the lower bound
    r_end, hist = adam(lambda r: (r, 5.0), 0.5, 8, 0.05, r_min=0.3, r_max=0.8)
    rs = [h["r"] for h in hist]
    assert rs[:5] == pytest.approx([0.5, 0.45, 0.40, 0.35, 0.30], abs=1e-9)
    assert all(0.3 <= r <= 0.8 for r in rs) and r_end == 0.3
    assert [h["iter"] for h in hist] == list(range(8))
    assert all(h["loss"] == h["r"] for h in hist)
    r_end, _ = adam(lambda r: (r, -5.0), 0.5, 10, 0.05, r_min=0.3, r_max=0.8)
    assert r_end == 0.8
