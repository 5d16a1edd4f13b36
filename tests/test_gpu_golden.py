"""The benchmark workloads on the HIP kernels against the golden fixtures.

The fixtures (tests/golden/*.json, written by tests/golden/make_golden.py) are
outputs of the CPU oracle at full length:
* 64x21x21 channel, 5000 BGK steps: u_x(y) profile (float64 and float32 oracle);
* 64x32x32 sphere, MRT + Bouzidi, 8000 steps: Cd of every step;
* 96x48x48 sphere, MRT + Bouzidi q(r), 2000 steps: dCd/dr of the mean Cd over
  the last 600 steps, by torch autograd of the float64 oracle.

Tolerances: float64 kernels agree with the float64 oracle to rounding.  The
float32 kernels meet the bars of the workload table for the profile (1e-5
absolute) and dCd/dr (1e-3 relative).

The 64x32x32 sphere starts from rest in a state symmetric in y and z (uniform
populations, the sphere on the centre line), and its Cd swings between 0 and
22 in the first 2000 steps.  Per step, float32 leaves float64 by more than
1e-4 within the first 15-45 steps, in the kernels and in the oracle alike, so
no early window holds a 1e-4 bar.  Unperturbed, the wake's symmetry is
broken by rounding alone, and the tail mean of that one run is +2.3e-4
(kernels) or +8e-6 (oracle) from float64: it is held to 5e-4.  From starts
perturbed by 1e-8 (tests/_helpers.py perturbed_start) the float32 kernels'
tail mean is -6.1e-5 +- 2.1e-5 from float64 over 32 seeds (one at -1.03e-4),
the float32 oracle's +3.0e-5 +- 3.1e-5 over 11 (tests/golden/c2_sphere_cd_
64x32x32_perturbed_f32.json), and the float64 kernels' -2e-7 +- 1e-6: the
1e-4 bar is pinned on the mean over eight perturbed starts, each within 2e-4.
"""

import json
import os

import pytest
import torch

from lbm_jax_autodiff_amd import protocols as P

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"golden fixture {name} not generated")
    with open(path) as fh:
        return json.load(fh)


@pytest.mark.parametrize("storage,tol", [("f64", 1e-11), ("f32", 1e-5)])
def test_c1_channel_profile(dev, storage, tol):
    g = _load("c1_channel_64x21x21.json")
    ref = torch.tensor(g["profiles"]["f64"], dtype=torch.float64)
    res = P.channel_flow(nx=64, ny=21, nz=21, tau=0.8, u_in=0.05, n_steps=5000, storage=storage)
    prof = torch.tensor(res["profile"], dtype=torch.float64)
    err = float((prof - ref).abs().max())
    print(f"c1 {storage}: max |u_x - u_x,oracle f64| = {err:.3e}")
    # the protocol rounds its profile to 8 decimals
    assert err <= tol + 5e-9, err


def _tail_mean(cd, frac=0.3):
    n = len(cd)
    k = max(1, int(round(frac * n)))
    return sum(cd[n - k:]) / k


@pytest.mark.parametrize("storage,tol", [("f64", 1e-9), ("f32", 5e-4)])
def test_c2_time_averaged_cd(dev, storage, tol):
    g = _load("c2_sphere_cd_64x32x32.json")
    nx, ny, nz = 64, 32, 32
    res, cd = P.sphere_cd(nx, ny, nz, g["steps"], P.world_center(nx, ny, nz, (16, 16, 16)), 0.5, storage=storage)
    assert res["area"] == g["f64"]["area"]
    ref = _tail_mean(g["f64"]["cd"])
    rel = abs(res["cd_tail30_mean"] - ref) / abs(ref)
    print(f"c2 {storage}: Cd tail mean {res['cd_tail30_mean']:.9f} oracle f64 {ref:.9f} rel {rel:.3e}")
    assert rel <= tol, rel
    if storage == "f32" and "f32" in g:
        ref32 = _tail_mean(g["f32"]["cd"])
        rel32 = abs(res["cd_tail30_mean"] - ref32) / abs(ref32)
        print(f"c2 f32: oracle f32 {ref32:.9f} rel {rel32:.3e}")
        assert rel32 <= tol, rel32
    if storage == "f64":
        # the whole series, step by step
        ref_series = torch.tensor(g["f64"]["cd"], dtype=torch.float64)
        got = cd.detach().cpu().to(torch.float64)
        assert float(((got - ref_series).abs() / ref_series.abs().clamp(min=1e-3)).max()) <= 1e-8


@pytest.mark.parametrize("storage", ["f32", "f64"])
def test_c2_time_averaged_cd_perturbed_ensemble(dev, storage):
    """Tail-mean Cd from eight initial states perturbed by 1e-8 (seeds 1-8):
    float32 mean within 1e-4 of the float64 oracle and each run within 2e-4;
    float64 each within 5e-6."""
    from _helpers import perturbed_start
    from lbm_jax_autodiff_amd.adjoint import drag_series
    g = _load("c2_sphere_cd_64x32x32.json")
    ref = _tail_mean(g["f64"]["cd"])
    nx, ny, nz = 64, 32, 32
    lat, _, _, area = P.sphere_lattice(nx, ny, nz, P.world_center(nx, ny, nz, (16, 16, 16)), 0.5, storage=storage)
    dt = torch.float64 if storage == "f64" else torch.float32
    rels = []
    for seed in range(1, 9):
        _, _, F = lat.run(perturbed_start(nx, ny, nz, seed).to(dt).to(dev), g["steps"])
        cd = drag_series(F, 0.05, area).detach().cpu().to(torch.float64).tolist()
        rels.append(_tail_mean(cd) / ref - 1.0)
    mean = sum(rels) / len(rels)
    print(f"c2 {storage} perturbed: mean rel {mean:.3e}, runs {[f'{r:.2e}' for r in rels]}")
    if storage == "f32":
        assert abs(mean) <= 1e-4 and max(abs(r) for r in rels) <= 2e-4
    else:
        assert max(abs(r) for r in rels) <= 5e-6


@pytest.mark.parametrize("storage,tol", [("f64", 1e-8), ("f32", 1e-3)])
def test_c3_dcd_dr(dev, storage, tol):
    g = _load("c3_sphere_dcd_dr_96x48x48.json")
    nx, ny, nz = 96, 48, 48
    res = P.sphere_loss_and_grad(nx, ny, nz, g["steps"], P.world_center(nx, ny, nz, (24, 24, 24)), 0.5,
                                 avg_last=g["avg_last"], storage=storage)
    rel_l = abs(res["loss"] - g["loss"]) / abs(g["loss"])
    rel_g = abs(res["dloss_dr_world"] - g["dloss_dr_world"]) / abs(g["dloss_dr_world"])
    print(f"c3 {storage}: loss {res['loss']:.9f} (oracle {g['loss']:.9f}, rel {rel_l:.2e}); dCd/dr "
          f"{res['dloss_dr_world']:.9f} (oracle autograd f64 {g['dloss_dr_world']:.9f}, rel {rel_g:.2e})")
    assert rel_l <= tol and rel_g <= tol, (rel_l, rel_g)


def _c3(radius, storage):
    nx, ny, nz = 96, 48, 48
    torch.cuda.empty_cache()
    return P.sphere_loss_and_grad(nx, ny, nz, 2000, P.world_center(nx, ny, nz, (24, 24, 24)), radius, avg_last=600,
                                  storage=storage)


def test_c3_adjoint_matches_central_differences(dev):
    """The float64 adjoint against central differences of the same kernels'
    loss: second order in h (2.1e-4 relative at h = 1e-3, 2e-8 at 1e-5), so
    the gradient carries no missing or extra term."""
    ad = _c3(0.5, "f64")["dloss_dr_world"]
    fd = {}
    for h in (1e-3, 1e-5):
        fd[h] = (_c3(0.5 + h, "f64")["loss"] - _c3(0.5 - h, "f64")["loss"]) / (2.0 * h)
    e3, e5 = abs(fd[1e-3] / ad - 1.0), abs(fd[1e-5] / ad - 1.0)
    print(f"c3 f64 adjoint {ad:.12f}; FD h=1e-3 {fd[1e-3]:.12f} ({e3:.2e}), h=1e-5 {fd[1e-5]:.12f} ({e5:.2e})")
    assert e5 < 1e-6
    assert 1e-5 < e3 < 1e-3  # the O(h^2) truncation error is visible at h = 1e-3


def test_c3_adam_twenty_iterations(dev):
    """The C3 optimisation: 20 Adam iterations on r (lr 0.05, clipped to
    [0.3, 0.8]) from r = 0.5 with float32 gradients.  The gradient stays
    positive, so r falls by about lr per iteration to the lower bound, which
    it reaches at iteration 4 and keeps; the loss at r = 0.5 is the C3 loss."""
    seen = []

    def lg(r):
        out = _c3(r, "f32")
        seen.append((r, out["loss"], out["dloss_dr_world"]))
        return out["loss"], out["dloss_dr_world"]

    r, hist = P.adam(lg, 0.5, 20, 0.05)
    print("c3 adam:", [(round(h["r"], 6), round(h["loss"], 6), round(h["grad"], 4)) for h in hist])
    assert len(hist) == 20 and len(seen) == 20
    assert all(h["grad"] > 0.0 for h in hist)
    assert abs(hist[0]["loss"] - 1.2723378) < 1e-6
    # Adam's first step is lr times the gradient's sign
    assert abs(hist[1]["r"] - 0.45) < 1e-6
    first = min(t for t, h in enumerate(hist) if h["r"] == 0.3)
    assert first in (4, 5)
    assert all(0.04 < hist[t]["r"] - hist[t + 1]["r"] < 0.06 for t in range(first - 1))
    assert all(h["r"] == 0.3 for h in hist[first:]) and r == 0.3
    assert hist[-1]["loss"] > hist[0]["loss"]
