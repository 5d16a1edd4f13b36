"""Layout plumbing between the reference's (19, nx, ny, nz) indexing and the
library's lattice storage (x fastest: a contiguous (19, nz, ny, nx) tensor).

A tensor produced by this package is a permuted view of lattice storage, so
passing it back in costs no copy.  Any other (19, nx, ny, nz) tensor is copied
into lattice storage once.
"""

from __future__ import annotations

import torch

from . import _lib


def dtype_tag(dtype: torch.dtype) -> int:
    if dtype == torch.float32:
        return _lib.LBM_F32
    if dtype == torch.float64:
        return _lib.LBM_F64
    raise TypeError(f"unsupported dtype {dtype}: the per-operator surface computes in float32 or float64")


def require_cuda(t: torch.Tensor, name: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor on a HIP device, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on a HIP device (got device {t.device}); "
                         "this package has no CPU path")


def require_shape(t: torch.Tensor, name: str, shape, device=None) -> None:
    """t must be a device tensor of exactly ``shape`` (and on ``device`` when
    given): the kernels index these buffers by the grid's sizes, so a host
    tensor or a short one must fail here, not as a GPU memory fault."""
    require_cuda(t, name)
    if tuple(t.shape) != tuple(int(s) for s in shape):
        raise ValueError(f"{name} must have shape {tuple(int(s) for s in shape)}, got {tuple(t.shape)}")
    if device is not None and not same_device(t.device, device):
        raise ValueError(f"{name} lives on {t.device}, expected {device}")


def same_device(a, b) -> bool:
    """Device equality with an index-less "cuda" read as the current device."""
    a, b = torch.device(a), torch.device(b)
    if a.type != b.type:
        return False
    if a.type != "cuda":
        return True
    ia = a.index if a.index is not None else torch.cuda.current_device()
    ib = b.index if b.index is not None else torch.cuda.current_device()
    return ia == ib


def dims_of_pop(f: torch.Tensor):
    if f.dim() != 4 or f.shape[0] != 19:
        raise ValueError(f"expected populations of shape (19, nx, ny, nz), got {tuple(f.shape)}")
    return int(f.shape[1]), int(f.shape[2]), int(f.shape[3])


def pop_storage(f: torch.Tensor) -> torch.Tensor:
    """(19, nx, ny, nz) -> contiguous (19, nz, ny, nx)."""
    return f.permute(0, 3, 2, 1).contiguous()


def pop_view(s: torch.Tensor) -> torch.Tensor:
    return s.permute(0, 3, 2, 1)


def scalar_storage(a: torch.Tensor) -> torch.Tensor:
    """(nx, ny, nz) -> contiguous (nz, ny, nx)."""
    return a.permute(2, 1, 0).contiguous()


def scalar_view(s: torch.Tensor) -> torch.Tensor:
    return s.permute(2, 1, 0)


def vec_storage(u: torch.Tensor) -> torch.Tensor:
    """(3, nx, ny, nz) -> contiguous (3, nz, ny, nx)."""
    return u.permute(0, 3, 2, 1).contiguous()


def vec_view(s: torch.Tensor) -> torch.Tensor:
    return s.permute(0, 3, 2, 1)


def new_pop(nx, ny, nz, dtype, device) -> torch.Tensor:
    return torch.empty((19, nz, ny, nx), dtype=dtype, device=device)


def new_scalar(nx, ny, nz, dtype, device) -> torch.Tensor:
    return torch.empty((nz, ny, nx), dtype=dtype, device=device)


def as_field(v, like: torch.Tensor, shape3) -> torch.Tensor:
    """Scalar (Python number or 0-d tensor) or (nx, ny, nz) field -> lattice-storage field."""
    nx, ny, nz = shape3
    if isinstance(v, torch.Tensor) and v.dim() > 0:
        require_shape(v, "field", (nx, ny, nz), like.device)
        return scalar_storage(v.to(like.dtype))
    val = float(v)
    return torch.full((nz, ny, nx), val, dtype=like.dtype, device=like.device)


def scalar_float(v) -> float:
    if isinstance(v, torch.Tensor):
        if v.numel() != 1:
            raise ValueError("expected a scalar")
        return float(v.detach().reshape(()).item())
    return float(v)
