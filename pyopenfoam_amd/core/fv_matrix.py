"""FvMatrix: LduMatrix with a source vector (reference ``core/fv_matrix.py``).

``source`` (:120-127), ``relaxation_factor`` (:129-132),
``add_boundary_contribution(bc, field)`` (:138-172: the boundary condition's
``matrix_contributions(field, n_cells, diag=, source=)`` returns the new
diagonal and source), ``add_explicit_source`` (:174-186), ``relax(field_old,
alpha)`` (:192-235: diag / alpha and source + (1 - alpha) diag' field_old, the
``pfm_relax`` kernel), ``set_reference`` (:241-259: diag[c] += max(|diag[c]|,
1) * 1e10 and the matching source), ``solve(solver, x0, tolerance,
max_iter)`` (:265-290) and ``residual(x) = b - A x`` (the ``pfm_residual``
kernel).
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
        self._field_old: torch.Tensor | None = None
        self._relaxation_factor: float = 1.0

    @property
    def source(self) -> torch.Tensor:
        return self._source

    @source.setter
    def source(self, value: torch.Tensor) -> None:
        self._source = value.to(device=self._device, dtype=self._dtype)

    @property
    def relaxation_factor(self) -> float:
        return self._relaxation_factor

    def add_boundary_contribution(self, bc, field: torch.Tensor | None = None) -> None:
        """The boundary condition adds its implicit (diagonal) and explicit
        (source) terms through ``bc.matrix_contributions``."""
        if field is None:
            field = torch.zeros(self._n_cells, device=self._device, dtype=self._dtype)
        else:
            field = field.to(device=self._device, dtype=self._dtype)
        diag, source = bc.matrix_contributions(field, self._n_cells, diag=self._diag,
                                               source=self._source)
        self._diag = diag
        self._source = source
        self.invalidate_cache()

    def add_explicit_source(self, values: torch.Tensor) -> None:
        if values.dim() == 0:
            self._source = self._source + values
        else:
            self._source = self._source + values.to(device=self._device, dtype=self._dtype)

    def relax(self, field_old: torch.Tensor, under_relaxation_factor: float = 1.0) -> None:
        """Implicit under-relaxation: diag' = diag / alpha,
        source' = source + (1 - alpha) * diag' * field_old (alpha < 1 only)."""
        assert_floating(field_old, "field_old")
        alpha = float(under_relaxation_factor)
        if alpha <= 0.0 or alpha > 1.0:
            raise ValueError(f"under_relaxation_factor must be in (0, 1], got {alpha}")
        self._field_old = field_old.to(device=self._device, dtype=self._dtype)
        self._relaxation_factor = alpha
        if alpha < 1.0:
            self._check_compute()
            if self._field_old.shape[0] != self._n_cells:
                raise ValueError(f"field_old has {self._field_old.shape[0]} elements, "
                                 f"expected {self._n_cells}")
            d = self._diag.detach().contiguous()
            src = self._source.detach().contiguous()
            phi = self._field_old.detach().contiguous()
            d_new, s_new = torch.empty_like(d), torch.empty_like(src)
            if self._n_cells:
                _lib.check(_lib.lib().pfm_relax(self._n_cells, alpha, _lib.ptr(d), _lib.ptr(src),
                                                _lib.ptr(phi), _lib.ptr(d_new), _lib.ptr(s_new),
                                                current_stream(self._device)), "pfm_relax")
            self._diag, self._source = d_new, s_new
            self.invalidate_cache()

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
