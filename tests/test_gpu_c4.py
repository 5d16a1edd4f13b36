"""The 256x128x128 precision study: dCd/dr of the sphere (D = 32, Re = 100,
MRT + Bouzidi q(r), pre guard) over 28,000 steps with checkpoints every 100
steps, the loss being Cd averaged over the last 30% of the steps.  Float32
and the 16-bit storages (deviations from the free-stream equilibrium,
stochastic rounding, half cotangents) against float64 storage of the same
kernels: Cd and dCd/dr within 1e-3 relative for the 16-bit formats, and the
float16 pipeline at least 1.6 times faster than the float32 one."""

import pytest
import torch

from lbm_jax_autodiff_amd.protocols import sphere_loss_and_grad

pytestmark = pytest.mark.gpu

NX, NY, NZ = 256, 128, 128
STEPS = 28000


def _study(storage, steps=STEPS):
    torch.cuda.empty_cache()
    return sphere_loss_and_grad(NX, NY, NZ, steps, (0.0, 0.0, 0.0), 0.5, storage=storage, checkpoint=100,
                                avg_last=int(0.3 * steps), guard="pre")


@pytest.fixture(scope="module")
def f64_reference(dev):
    return _study("f64")


def _rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.parametrize("storage,tol", [("f32", 1e-4), ("f16", 1e-3), ("bf16", 1e-3)])
def test_28000_step_gradient_against_float64(dev, f64_reference, storage, tol):
    r = _study(storage)
    ref = f64_reference
    e_cd, e_g = _rel(r["loss"], ref["loss"]), _rel(r["dloss_dr_world"], ref["dloss_dr_world"])
    print(f"{storage}: Cd {r['loss']:.9f} ({e_cd:.1e}), dCd/dr {r['dloss_dr_world']:.9f} ({e_g:.1e}), "
          f"{r['fwd_adj_s']:.2f} s")
    assert e_cd <= tol and e_g <= tol


def test_28000_step_float16_pipeline_faster_than_float32(dev):
    """Wall time of the whole forward + adjoint, each after a 200-step run of
    the same pipeline (allocator and plans warm)."""
    t = {}
    for st in ("f32", "f16"):
        _study(st, 200)
        t[st] = _study(st)["fwd_adj_s"]
    print(f"float32 {t['f32']:.2f} s, float16 {t['f16']:.2f} s: {t['f32'] / t['f16']:.2f}x")
    assert t["f32"] / t["f16"] >= 1.6
