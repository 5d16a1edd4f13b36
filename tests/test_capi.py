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
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|size_t)\s+(lbm_\w+)\s*\(", text, flags=re.M)))


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


def _c_offsets(struct, fields):
    """offsetof/sizeof of a header struct, from a C program compiled with gcc."""
    import shutil
    import subprocess
    import tempfile
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "lbm_hip.h"', 'int main(void) {']
    for f in fields:
        lines.append(f'  printf("%zu\\n", offsetof({struct}, {f}));')
    lines.append(f'  printf("%zu\\n", sizeof({struct}));')
    lines.append('  return 0; }')
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "layout.c")
        exe = os.path.join(td, "layout")
        with open(src, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


@pytest.mark.parametrize("cls,struct", [("StepDesc", "lbm_step_desc"), ("AdjointDesc", "lbm_adjoint_desc")])
def test_desc_layout_matches_header(cls, struct):
    """Every field of the ctypes mirrors sits at the C compiler's offset."""
    import lbm_jax_autodiff_amd._lib as L
    import lbm_jax_autodiff_amd.adjoint as A
    d = getattr(L, cls, None) or getattr(A, cls)
    names = [f[0] for f in d._fields_]
    offs = _c_offsets(struct, names)
    assert offs[:-1] == [getattr(d, n).offset for n in names]
    assert offs[-1] == ctypes.sizeof(d)


def test_no_cpu_fallback_without_device():
    """Operators refuse CPU tensors instead of computing on the host."""
    import torch
    import lbm_jax_autodiff_amd as lbm
    with pytest.raises(ValueError):
        lbm.compute_density(torch.zeros((19, 2, 2, 2)))
