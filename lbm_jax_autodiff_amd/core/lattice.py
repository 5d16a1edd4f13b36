"""D3Q19 lattice tables as host-side torch tensors (jax_lbm/core/lattice.py:29-129).

The ordering, weights, opposite map and MRT tables are part of the data
layout the kernels share with the reference, so they are identical to the
reference's.  The kernels hold the same tables as compile-time constants
(csrc/lbm_common.hpp).
"""
import torch

D3Q19_VELOCITIES = torch.tensor([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], dtype=torch.int32)
D3Q19_WEIGHTS = torch.tensor([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12, dtype=torch.float32)
D3Q19_OPPOSITE = torch.tensor([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15], dtype=torch.int32)
CS2 = 1.0 / 3.0
INV_CS2 = 3.0
INV_CS4 = 9.0
Q = 19
MRT_ROW_NORMS = torch.tensor([19.0, 2394.0, 252.0, 10.0, 40.0, 10.0, 40.0, 10.0, 40.0, 36.0, 72.0, 12.0, 24.0,
                              4.0, 4.0, 4.0, 8.0, 8.0, 8.0], dtype=torch.float32)
MRT_RELAXATION_RATES = torch.tensor([0.0, 1.19, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 1.2, 0.0, 1.4, 0.0, 1.4, 0.0, 0.0, 0.0,
                                     1.98, 1.98, 1.98], dtype=torch.float32)
MRT_STRESS_INDICES = torch.tensor([9, 11, 13, 14, 15], dtype=torch.int32)

__all__ = ["D3Q19_VELOCITIES", "D3Q19_WEIGHTS", "D3Q19_OPPOSITE", "CS2", "INV_CS2", "INV_CS4", "Q",
           "MRT_ROW_NORMS", "MRT_RELAXATION_RATES", "MRT_STRESS_INDICES"]
