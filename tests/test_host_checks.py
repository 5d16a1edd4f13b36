"""Input checks of the host layer (no GPU needed): host tensors and wrong
shapes must raise a Python error before any kernel sees a pointer."""

import pytest
import torch

import lbm_jax_autodiff_amd as lbm
from lbm_jax_autodiff_amd import _field as F
from lbm_jax_autodiff_amd.core import ops
from lbm_jax_autodiff_amd.engine import Lattice

NX, NY, NZ = 8, 4, 4


@pytest.mark.parametrize("arg", ["solid", "porosity", "q"])
def test_lattice_rejects_host_tensors(arg):
    t = {"solid": torch.zeros((NX, NY, NZ), dtype=torch.int32), "porosity": torch.zeros((NX, NY, NZ)),
         "q": -torch.ones((19, NX, NY, NZ))}[arg]
    with pytest.raises(ValueError, match="HIP device"):
        Lattice(NX, NY, NZ, tau=0.6, device="cuda:0", **{arg: t})


def test_operators_reject_host_tensors():
    f = torch.zeros((19, NX, NY, NZ))
    with pytest.raises(ValueError, match="HIP device"):
        ops.hard_bounce_back(f, torch.zeros((NX, NY, NZ), dtype=torch.int32))
    with pytest.raises(ValueError, match="HIP device"):
        lbm.step(lbm.LBMState(f, torch.zeros((NX, NY, NZ), dtype=torch.int32), -torch.ones_like(f)),
                 lbm.LBMParams(0.6, (0.05, 0.0, 0.0)))
    with pytest.raises(TypeError):
        F.require_shape([1.0], "x", (1,))


def test_require_shape_messages():
    with pytest.raises(ValueError, match="HIP device"):
        F.require_shape(torch.zeros(3), "solid", (3,))
    assert F.same_device("cpu", torch.device("cpu"))
    assert not F.same_device("cpu", "cuda:0")


def test_lattice_rejects_unknown_rounding():
    with pytest.raises(ValueError, match="rounding"):
        Lattice(NX, NY, NZ, tau=0.6, storage="f16", rounding="up", device="cuda:0")


def test_rollout_rejects_unknown_rounding_before_any_launch():
    from lbm_jax_autodiff_amd import adjoint as A
    f = torch.zeros((19, NX, NY, NZ))
    with pytest.raises(ValueError, match="rounding"):
        A.rollout(f, 3, tau=0.6, storage="f16", rounding="up")
