"""The C-ABI library loads (without a GPU) and exports every symbol the header declares."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t)\s+(lbm_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("lbm_step", "lbm_stream", "lbm_mrt", "lbm_bgk", "lbm_momentum_exchange", "lbm_sphere_geometry",
                 "lbm_links_build", "lbm_force_reduce"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import lbm_jax_autodiff_amd._lib as L
    if not os.path.exists(L.LIB_PATH):
        pytest.skip("liblbm_hip.so not built")
    lib = L.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.lbm_version() >= 10000


def test_binding_signatures_cover_header():
    import lbm_jax_autodiff_amd._lib as L
    bound = set(L.exported_names())
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, missing


def test_step_desc_layout_matches_header():
    """ctypes mirror of lbm_step_desc: 10 int32, 2 doubles, u_in[3], rho_out,
    u_ref[3], offset[19], 4 pointers."""
    import lbm_jax_autodiff_amd._lib as L
    d = L.StepDesc
    assert d.tau.offset == 40
    assert d.offset.offset == 40 + 8 * (2 + 3 + 1 + 3)
    assert ctypes.sizeof(d) == d.link_q.offset + 8


def test_no_cpu_fallback_without_device():
    """Operators refuse CPU tensors instead of computing on the host."""
    import torch
    import lbm_jax_autodiff_amd as lbm
    with pytest.raises(ValueError):
        lbm.compute_density(torch.zeros((19, 2, 2, 2)))
