"""jax_lbm.core equivalents (jax_lbm/core/*.py) on liblbm_hip.so."""
from .lattice import *  # noqa: F401,F403
from .ops import (  # noqa: F401
    bgk, compute_density, compute_velocity, drag_coefficient, equilibrium, hard_bounce_back, momentum_exchange,
    momentum_exchange_decomposed, mrt, projected_area, smagorinsky_tau, soft_bounce_back, stability_guard, stream,
    zou_he_inlet, zou_he_outlet,
)
