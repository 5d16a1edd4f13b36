"""jax_lbm/core/collision.py surface."""
from .ops import bgk, mrt, smagorinsky_tau  # noqa: F401
