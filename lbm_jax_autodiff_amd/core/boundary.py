"""jax_lbm/core/boundary.py surface."""
from .ops import zou_he_inlet, zou_he_outlet  # noqa: F401
