"""Device and dtype defaults.

The reference takes its device from a global ``TensorConfig``
(``core/device.py:139-300``) and computes in float64 with int64 indices
(``core/dtype.py:38,47``).  Here compute always runs on a HIP device through
libpfm; the default device is the current HIP device and the default dtype is
float64.  Indices are stored int64 on the Python side (as the reference) and
handed to the native library as int32.
"""

from __future__ import annotations

import torch

CFD_DTYPE = torch.float64
INDEX_DTYPE = torch.int64

__all__ = ["CFD_DTYPE", "INDEX_DTYPE", "get_device", "get_default_dtype", "require_gpu",
           "assert_floating", "current_stream"]


def get_device() -> torch.device:
    if torch.cuda.is_available():
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device("cuda")


def get_default_dtype() -> torch.dtype:
    return CFD_DTYPE


def require_gpu(*tensors: torch.Tensor) -> None:
    """Raise unless every tensor lives on a HIP device (there is no CPU path)."""
    for t in tensors:
        if t is not None and t.device.type != "cuda":
            raise RuntimeError(
                "pyopenfoam_amd computes on the GPU only (HIP kernels in libpfm); "
                f"got a tensor on {t.device}")


def assert_floating(t, name: str) -> None:
    """TypeError for non-floating input, as ``core/dtype.py:581-590``."""
    if not isinstance(t, torch.Tensor) or not t.is_floating_point():
        raise TypeError(f"{name} must be a floating-point tensor, got "
                        f"{getattr(t, 'dtype', type(t))}")


def current_stream(device: torch.device) -> int:
    """Raw hipStream_t of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream
