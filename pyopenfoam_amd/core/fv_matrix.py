"""FvMatrix: LduMatrix with a source vector (reference ``core/fv_matrix.py``).

The subset the pressure-equation path uses: ``source`` (:101-107),
``set_reference`` (:241-259: diag[c] += max(|diag[c]|, 1) * 1e10 and the
matching source), ``solve(solver, x0, tolerance, max_iter)`` (:265-290) and
``residual(x) = b - A x`` (computed by the ``pfm_residual`` kernel).
"""

from __future__ import annotations

import torch

from .. import _lib
from .device import assert_floating, current_stream
from .ldu_matrix import LduMatrix

__all__ = ["FvMatrix"]


class FvMatrix(LduMatrix):
    def __init__(self, n_cells: int, owner: torch.Tensor, neighbour: torch.Tensor, *,
                 device: torch.device | str | None = None, dtype: torch.dtype | None = None) -> None:
        super().__init__(n_cells, owner, neighbour, device=device, dtype=dtype)
        self._source = torch.zeros(n_cells, device=self._device, dtype=self._dtype)

    @property
    def source(self) -> torch.Tensor:
        return self._source

    @source.setter
    def source(self, value: torch.Tensor) -> None:
        self._source = value.to(device=self._device, dtype=self._dtype)

    def set_reference(self, cell_index: int, value: float = 0.0) -> None:
        if cell_index < 0 or cell_index >= self._n_cells:
            raise ValueError(f"cell_index {cell_index} out of range [0, {self._n_cells})")
        large = self._diag[cell_index].abs().clamp(min=1.0) * 1e10
        self._diag[cell_index] += large
        self._source[cell_index] += large * value
        self.invalidate_cache()

    def solve(self, solver, x0: torch.Tensor, tolerance: float = 1e-6,
              max_iter: int = 1000) -> tuple[torch.Tensor, int, float]:
        assert_floating(x0, "x0")
        x0 = x0.to(device=self._device, dtype=self._dtype)
        return solver(self, self._source, x0, tolerance, max_iter)

    def residual(self, x: torch.Tensor) -> torch.Tensor:
        x = self._check_x(x)
        self._check_compute()
        r = torch.empty_like(x)
        if self._n_cells:
            d = self._diag.detach().contiguous()
            src = self._source.detach().contiguous()
            _lib.check(_lib.lib().pfm_residual(self.addressing.handle, _lib.ptr(d),
                                               _lib.ptr(self.packed(False)), _lib.ptr(x.detach()),
                                               _lib.ptr(src), _lib.ptr(r),
                                               current_stream(self._device)),
                       "pfm_residual")
        return r
