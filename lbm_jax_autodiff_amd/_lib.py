"""ctypes binding of liblbm_hip.so (the C ABI declared in include/lbm_hip.h).

The library is built in-tree by ``lbm_jax_autodiff_amd.build_lib``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbm_hip.so")

LBM_F32, LBM_F64, LBM_F16DEV, LBM_BF16DEV = 0, 1, 2, 3
LBM_COLL_BGK, LBM_COLL_MRT = 0, 1
LBM_BOUNCE_NONE, LBM_BOUNCE_HARD, LBM_BOUNCE_SOFT = 0, 1, 2
LBM_GUARD_NONE, LBM_GUARD_POST, LBM_GUARD_PRE = 0, 1, 2
LBM_MODE_FULL, LBM_MODE_COLLIDE, LBM_MODE_STREAM = 0, 1, 2
LBM_ROUND_NEAREST, LBM_ROUND_STOCHASTIC = 0, 1
LBM_NOALLOC_PARTIALS = 1 << 20  # doubles of the no-allocation entry point's partials array

_ERRORS = {1001: "LBM_E_ARG (bad size, pointer or flag)", 1002: "LBM_E_DTYPE (dtype not supported)",
           1003: "LBM_E_LAUNCH (kernel launch failed)"}

c_int, c_double, c_void_p, c_size_t = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
c_i32p = ctypes.POINTER(ctypes.c_int32)


class StepDesc(ctypes.Structure):
    """Mirror of ``lbm_step_desc`` (include/lbm_hip.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("storage", ctypes.c_int32), ("collision", ctypes.c_int32),
        ("smagorinsky", ctypes.c_int32), ("bounce", ctypes.c_int32),
        ("guard", ctypes.c_int32), ("periodic_yz", ctypes.c_int32),
        ("n_links", ctypes.c_int32),
        ("tau", ctypes.c_double), ("smag_cs", ctypes.c_double),
        ("u_in", ctypes.c_double * 3), ("rho_out", ctypes.c_double),
        ("u_ref", ctypes.c_double * 3), ("offset", ctypes.c_double * 19),
        ("flags", ctypes.c_void_p), ("porosity", ctypes.c_void_p),
        ("link_cell", ctypes.c_void_p), ("link_q", ctypes.c_void_p),
        ("cells_per_thread", ctypes.c_int32), ("pop_stride", ctypes.c_int64),
        ("row_stride", ctypes.c_int64), ("rounding", ctypes.c_int32), ("seed", ctypes.c_uint32),
        ("step", ctypes.c_int64),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "lbm_version": (c_int, []),
    "lbm_device_check": (c_int, []),
    "lbm_equilibrium": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_density": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_velocity": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_bgk": (c_int, [c_int, c_void_p, c_double, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_mrt": (c_int, [c_int, c_void_p, c_double, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_smagorinsky_tau": (c_int, [c_int, c_void_p, c_double, c_double, c_void_p, c_int, c_int, c_int,
                                    c_void_p]),
    "lbm_soft_bounce_back": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_hard_bounce": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_stability_guard": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_stability_guard_vjp": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_stream": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_zou_he_inlet": (c_int, [c_int, c_void_p, c_double, c_double, c_double, c_int, c_double, c_int, c_int,
                                 c_int, c_void_p]),
    "lbm_zou_he_outlet": (c_int, [c_int, c_void_p, c_double, c_int, c_int, c_int, c_void_p]),
    "lbm_reduce_workspace_bytes": (c_size_t, [c_int, c_int, c_int]),
    "lbm_momentum_exchange": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int,
                                      c_int, c_int, c_void_p]),
    "lbm_projected_area": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_drag_coefficient": (c_int, [c_void_p, c_int, c_double, c_void_p, c_double, c_void_p, c_void_p]),
    "lbm_sphere_geometry": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_double,
                                    c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                    c_int, c_int, c_void_p]),
    "lbm_dot": (c_int, [c_int, c_void_p, c_int, c_void_p, c_size_t, c_void_p, c_void_p, c_void_p]),
    "lbm_sphere_sdf": (c_int, [c_int, c_int, c_double, c_double, c_double, c_double, c_void_p, c_int, c_int,
                               c_int, c_void_p]),
    "lbm_porosity_from_sdf": (c_int, [c_int, c_void_p, c_double, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_q_from_solid": (c_int, [c_int, c_void_p, c_double, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_links_count": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_links_build": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_step_force_slots": (c_int, [ctypes.POINTER(StepDesc)]),
    "lbm_state_stride": (ctypes.c_int64, [ctypes.POINTER(StepDesc)]),
    "lbm_step_chain": (c_int, [ctypes.POINTER(StepDesc), c_int, c_void_p, c_void_p, ctypes.c_int64, c_void_p]),
    "lbm_adjoint_chain": (c_int, [ctypes.POINTER(StepDesc), c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p]),
    "lbm_step": (c_int, [ctypes.POINTER(StepDesc), c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "lbm_force_reduce": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "lbm_step_chain_forces_noalloc": (c_int, [ctypes.POINTER(StepDesc), c_int, c_void_p, c_void_p, c_void_p]),
    "lbm_convert_state": (c_int, [ctypes.POINTER(StepDesc), c_int, c_void_p, c_int, c_void_p, c_void_p]),
    "lbm_step_adjoint": (c_int, [ctypes.POINTER(StepDesc), ctypes.c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p]),
    "lbm_adjoint_scatter": (c_int, [ctypes.POINTER(StepDesc), ctypes.c_void_p, c_void_p, c_void_p, c_void_p]),
    "lbm_adjoint_link_finish": (c_int, [ctypes.POINTER(StepDesc), ctypes.c_void_p, c_void_p, c_void_p, c_void_p]),
    "lbm_equilibrium_vjp": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                    c_void_p]),
    "lbm_density_vjp": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_velocity_vjp": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                                 c_void_p]),
    "lbm_collide_vjp": (c_int, [c_int, c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                c_int, c_int, c_void_p]),
    "lbm_smagorinsky_tau_vjp": (c_int, [c_int, c_void_p, c_double, c_double, c_void_p, c_void_p, c_void_p, c_int,
                                        c_int, c_int, c_void_p]),
    "lbm_soft_bounce_back_vjp": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                         c_int, c_void_p]),
    "lbm_hard_bounce_vjp": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_stream_vjp": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                               c_void_p]),
    "lbm_zou_he_inlet_vjp": (c_int, [c_int, c_void_p, c_double, c_double, c_double, c_int, c_double, c_void_p,
                                     c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_zou_he_outlet_vjp": (c_int, [c_int, c_double, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    "lbm_momentum_exchange_vjp": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                          c_int, c_int, c_void_p]),
    "lbm_drag_coefficient_vjp": (c_int, [c_void_p, c_void_p, c_int, c_double, c_void_p, c_double, c_void_p,
                                         c_void_p]),
    "lbm_sphere_geometry_vjp_workspace_bytes": (c_size_t, []),
    "lbm_sphere_geometry_vjp": (c_int, [c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_double,
                                        c_double, c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                        c_int, c_int, c_void_p]),
}

_lib = None


def exported_names():
    return sorted(_SIGS)


def load():
    """Load liblbm_hip.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m lbm_jax_autodiff_amd.build_lib` "
            "(the HIP library is required; there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    """Invoke a C-ABI entry point and raise on a non-zero status."""
    fn = getattr(load(), name)
    rc = fn(*args)
    if rc != 0:
        raise RuntimeError(f"{name} failed: {_ERRORS.get(rc, f'hipError {rc}')}")
    return rc


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def ptr_array(tensors):
    """Host array of device pointers (``void* const*`` arguments)."""
    return (c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])
