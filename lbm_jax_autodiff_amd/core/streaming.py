"""jax_lbm/core/streaming.py surface."""
from .ops import stream  # noqa: F401
