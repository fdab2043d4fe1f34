"""DIC and DILU preconditioners on the GPU (reference ``solvers/preconditioners.py``).

``DICPreconditioner`` (:62-201) and ``DILUPreconditioner`` (:204-337) factor
the diagonal with the reference's sequential recurrences, evaluated on the
device by libpfm in dependency order (``pfm_dic_factor``: one launch per cell
level; ``pfm_dilu_factor``: one persistent workgroup over face levels).  The
result is bit-identical to the sequential loops.  ``apply`` is the reference's
diagonal apply ``r / d`` (the Krylov solvers fuse it into their kernels);
``apply_full`` the triangular solve of the factored M (forward face sweep,
diagonal scaling, reverse face sweep; ``pfm_precond_apply_full``, level-scheduled
and bit-identical to the sequential sweeps).
"""

from __future__ import annotations

from typing import Protocol

import torch

from .. import _lib
from ..core.device import assert_floating, current_stream
from ..core.ldu_matrix import LduMatrix

__all__ = ["Preconditioner", "DICPreconditioner", "DILUPreconditioner", "make_preconditioner"]


class Preconditioner(Protocol):
    def apply(self, r: torch.Tensor) -> torch.Tensor:
        ...


class _DiagonalFactor:
    kind = ""

    def __init__(self, matrix: LduMatrix, *, transpose: bool = False) -> None:
        self._matrix = matrix
        self._device = matrix.device
        self._dtype = matrix.dtype
        self._transpose = bool(transpose)
        self._precond_diag = self._build()

    @property
    def diagonal(self) -> torch.Tensor:
        """The factored diagonal |d|.clamp(min=1e-30)."""
        return self._precond_diag

    def _build(self) -> torch.Tensor:
        raise NotImplementedError

    def apply(self, r: torch.Tensor) -> torch.Tensor:
        assert_floating(r, "r")
        r = r.to(device=self._device, dtype=self._dtype)
        return r / self._precond_diag

    def apply_full(self, r: torch.Tensor) -> torch.Tensor:
        """z = M^-1 r with M = (D + L) D^-1 (D + U) (reference :153-201, :290-337)."""
        assert_floating(r, "r")
        m = self._matrix
        r = r.to(device=self._device, dtype=self._dtype).contiguous()
        if r.shape[0] != m.n_cells:
            raise ValueError(f"r has {r.shape[0]} elements, expected {m.n_cells}")
        z = torch.empty_like(r)
        if m.n_cells == 0:
            return z
        lo, up = m.lower.detach().contiguous(), m.upper.detach().contiguous()
        if self._transpose:
            lo, up = up, lo
        _lib.check(_lib.lib().pfm_precond_apply_full(m.addressing.handle, int(self.kind == "DILU"),
                                                     _lib.ptr(self._precond_diag), _lib.ptr(lo),
                                                     _lib.ptr(up), _lib.ptr(r), _lib.ptr(z),
                                                     current_stream(self._device)),
                   "pfm_precond_apply_full")
        return z


class DICPreconditioner(_DiagonalFactor):
    kind = "DIC"

    def _build(self) -> torch.Tensor:
        m = self._matrix
        rd = torch.empty(m.n_cells, device=self._device, dtype=torch.float64)
        if m.n_cells == 0:
            return rd
        pk = m.packed(self._transpose)
        d = m.diag.detach().contiguous()
        _lib.check(_lib.lib().pfm_dic_factor(m.addressing.handle, _lib.ptr(d), _lib.ptr(pk),
                                             _lib.ptr(rd), current_stream(self._device)),
                   "pfm_dic_factor")
        return rd


class DILUPreconditioner(_DiagonalFactor):
    kind = "DILU"

    def _build(self) -> torch.Tensor:
        m = self._matrix
        rd = torch.empty(m.n_cells, device=self._device, dtype=torch.float64)
        if m.n_cells == 0:
            return rd
        d, lo, up = m.coefficients()
        _lib.check(_lib.lib().pfm_dilu_factor(m.addressing.handle, int(self._transpose),
                                              _lib.ptr(d), _lib.ptr(lo), _lib.ptr(up),
                                              _lib.ptr(rd), current_stream(self._device)),
                   "pfm_dilu_factor")
        return rd


def make_preconditioner(kind: str, matrix: LduMatrix, *, default_names: tuple[str, ...],
                        transpose: bool = False):
    k = kind.upper()
    if k == "DIC":
        return DICPreconditioner(matrix, transpose=transpose)
    if k == "DILU":
        return DILUPreconditioner(matrix, transpose=transpose)
    if k == "NONE":
        return None
    raise ValueError(f"Unknown preconditioner '{kind}'. Use {', '.join(repr(n) for n in default_names)}.")
