"""Smoothers (reference ``solvers/smoothers.py``) on the device.

``GaussSeidelSmoother(matrix, omega=1.0).smooth(x, b, n_iterations)``
(:63-149): n_iterations symmetric sweeps (forward over cells 0..n-1, then
backward), each cell updated with the latest values of its neighbours, row
sums over the cell's faces in face order.  ``pfm_gs_smooth`` runs the sweeps
level by level (cells of one dependency level are independent), which gives
the sequential sweeps' exact arithmetic.

``JacobiSmoother(matrix, omega=2/3).smooth(x, b, n_iterations)`` (:152-200):
x = x + (omega / max(|diag|, 1e-30)) * (b - A x), ``pfm_jacobi_smooth``.

``DICGSmoother(matrix, omega=1.0).smooth(x, b, n_iterations)`` (:203-260):
x = x + omega * DIC.apply_full(b - A x) per iteration.

All return a new tensor and leave ``x`` unchanged, as the reference does.
"""

from __future__ import annotations

import torch

from .. import _lib
from ..core.device import current_stream, require_gpu
from ..core.ldu_matrix import LduMatrix

__all__ = ["GaussSeidelSmoother", "JacobiSmoother", "DICGSmoother", "Smoother"]


class Smoother:
    """Protocol of the reference (``smoothers.py:45-60``): ``smooth(x, b, n)``."""

    def smooth(self, x: torch.Tensor, b: torch.Tensor, n_iterations: int) -> torch.Tensor:
        raise NotImplementedError


class _DeviceSmoother(Smoother):
    def __init__(self, matrix: LduMatrix, omega: float) -> None:
        self._matrix = matrix
        self._omega = float(omega)
        self._device = matrix.device
        self._dtype = matrix.dtype

    def _prep(self, x, b):
        require_gpu(self._matrix.diag)
        m = self._matrix
        xx = x.to(device=self._device, dtype=self._dtype).clone().contiguous()
        bb = b.to(device=self._device, dtype=self._dtype).contiguous()
        if xx.shape[0] != m.n_cells or bb.shape[0] != m.n_cells:
            raise ValueError(f"x and b must have {m.n_cells} elements")
        return xx, bb


class GaussSeidelSmoother(_DeviceSmoother):
    def __init__(self, matrix: LduMatrix, omega: float = 1.0) -> None:
        super().__init__(matrix, omega)

    def smooth(self, x: torch.Tensor, b: torch.Tensor, n_iterations: int) -> torch.Tensor:
        xx, bb = self._prep(x, b)
        m = self._matrix
        if n_iterations <= 0 or m.n_cells == 0:
            return xx
        d, pk = m.diag.detach().contiguous(), m.packed(False)
        _lib.check(_lib.lib().pfm_gs_smooth(m.addressing.handle, _lib.ptr(d),
                                            _lib.ptr(pk), _lib.ptr(bb), _lib.ptr(xx),
                                            int(n_iterations), self._omega,
                                            current_stream(self._device)), "pfm_gs_smooth")
        return xx


class JacobiSmoother(_DeviceSmoother):
    def __init__(self, matrix: LduMatrix, omega: float = 2.0 / 3.0) -> None:
        super().__init__(matrix, omega)

    def smooth(self, x: torch.Tensor, b: torch.Tensor, n_iterations: int) -> torch.Tensor:
        xx, bb = self._prep(x, b)
        m = self._matrix
        if n_iterations <= 0 or m.n_cells == 0:
            return xx
        work = torch.empty_like(xx)
        d, pk = m.diag.detach().contiguous(), m.packed(False)
        _lib.check(_lib.lib().pfm_jacobi_smooth(m.addressing.handle, _lib.ptr(d),
                                                _lib.ptr(pk), _lib.ptr(bb),
                                                _lib.ptr(xx), int(n_iterations), self._omega,
                                                _lib.ptr(work), current_stream(self._device)),
                   "pfm_jacobi_smooth")
        return xx


class DICGSmoother(_DeviceSmoother):
    def __init__(self, matrix: LduMatrix, omega: float = 1.0) -> None:
        super().__init__(matrix, omega)
        from .preconditioners import DICPreconditioner
        self._dic = DICPreconditioner(matrix)

    def smooth(self, x: torch.Tensor, b: torch.Tensor, n_iterations: int) -> torch.Tensor:
        xx, bb = self._prep(x, b)
        for _ in range(max(int(n_iterations), 0)):
            r = bb - self._matrix.Ax(xx)
            xx = xx + self._omega * self._dic.apply_full(r)
        return xx
