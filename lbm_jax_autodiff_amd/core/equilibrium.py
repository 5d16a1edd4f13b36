"""jax_lbm/core/equilibrium.py surface."""
from .ops import compute_density, compute_velocity, equilibrium  # noqa: F401
