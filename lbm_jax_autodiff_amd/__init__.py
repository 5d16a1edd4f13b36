"""lbm_jax_autodiff_amd: the jax_lbm per-timestep hot path on MI355X (gfx950).

Drop-in for the public surface re-exported at jax_lbm/__init__.py:23-60, with
torch tensors on a HIP device in place of jax arrays.  Every operator runs in
liblbm_hip.so (hand-written HIP kernels); importing this package does not load
the library, the first call does, and it raises if the library was not built.
"""
from .core.lattice import (  # noqa: F401
    CS2, D3Q19_OPPOSITE, D3Q19_VELOCITIES, D3Q19_WEIGHTS, MRT_RELAXATION_RATES, MRT_ROW_NORMS, MRT_STRESS_INDICES, Q,
)
from .core.ops import (  # noqa: F401
    bgk, compute_density, compute_velocity, drag_coefficient, equilibrium, momentum_exchange,
    momentum_exchange_decomposed, mrt, projected_area, smagorinsky_tau, soft_bounce_back, stream, zou_he_inlet,
    zou_he_outlet,
)
from .geometry import (  # noqa: F401
    compute_q_from_solid, create_sphere, porosity_from_sdf, signed_distance_sphere, smooth_sphere_geometry,
)
from .state import LBMParams, LBMState, simulate, step  # noqa: F401
from .engine import Lattice  # noqa: F401

__version__ = "0.1.0"
