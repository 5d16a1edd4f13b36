"""jax_lbm/core/forces.py surface."""
from .ops import drag_coefficient, momentum_exchange, momentum_exchange_decomposed, projected_area  # noqa: F401
