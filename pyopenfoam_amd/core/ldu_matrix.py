"""LDU matrix on the GPU: the reference's ``LduMatrix`` contract, HIP compute.

Storage contract of ``core/ldu_matrix.py:27-71``: ``diag`` (n_cells),
``lower[f] = A[owner[f], neighbour[f]]``, ``upper[f] = A[neighbour[f], owner[f]]``.
Setters cast to the matrix device/dtype and alias the caller's tensor when
nothing changes (``:143-166``); ``Ax``/``Ax_T`` check the input as ``:192-198``
(TypeError for non-float, ValueError for a wrong length) and run the
``pfm_spmv_ldu`` kernel.  There is no CPU compute path: matrices live on a HIP
device and every product is a libpfm kernel.

The packed (row-slice) copy of the coefficients used by the solvers is cached
per transpose flag and rebuilt when a coefficient tensor is replaced or
modified in place (tracked through the tensor version counters), the analogue
of the reference's CSR cache and ``invalidate_cache`` (``:309-343``).

``to_sparse_coo`` / ``to_sparse_csr`` / ``to_sparse_csr_cached`` (``:278-343``)
build torch sparse tensors of the same entries (format conversion, not
arithmetic); ``Ax_sparse`` / ``Ax_batched`` (``:345-381``) take (n,) or (n, k)
and run the multi-right-hand-side row-slice kernel ``pfm_spmv_multi``, whose
every column equals ``Ax`` of that column bit for bit (the reference's CSR
product rounds each row in column order instead).
"""

from __future__ import annotations

import torch

from .. import _lib
from .addressing import LduAddressing, addressing_for
from .device import (INDEX_DTYPE, assert_floating, current_stream, get_default_dtype, get_device,
                     require_gpu)

__all__ = ["LduMatrix"]


class LduMatrix:
    def __init__(self, n_cells: int, owner: torch.Tensor, neighbour: torch.Tensor, *,
                 device: torch.device | str | None = None, dtype: torch.dtype | None = None) -> None:
        self._device = torch.device(device) if device is not None else get_device()
        if self._device.type == "cuda" and self._device.index is None and torch.cuda.is_available():
            self._device = torch.device("cuda", torch.cuda.current_device())
        self._dtype = dtype or get_default_dtype()
        self._n_cells = int(n_cells)
        self._n_internal_faces = int(neighbour.shape[0])
        self._owner_src = owner
        self._neighbour_src = neighbour
        self._owner = owner.to(device=self._device, dtype=INDEX_DTYPE)
        self._neighbour = neighbour.to(device=self._device, dtype=INDEX_DTYPE)
        self._diag = torch.zeros(n_cells, device=self._device, dtype=self._dtype)
        self._lower = torch.zeros(self._n_internal_faces, device=self._device, dtype=self._dtype)
        self._upper = torch.zeros(self._n_internal_faces, device=self._device, dtype=self._dtype)
        self._addr: LduAddressing | None = None
        self._packed: dict = {}
        self._csr_cache: dict = {}

    # ------------------------------------------------------------ properties
    @property
    def n_cells(self) -> int:
        return self._n_cells

    @property
    def n_internal_faces(self) -> int:
        return self._n_internal_faces

    @property
    def device(self) -> torch.device:
        return self._device

    @property
    def dtype(self) -> torch.dtype:
        return self._dtype

    @property
    def owner(self) -> torch.Tensor:
        return self._owner

    @property
    def neighbour(self) -> torch.Tensor:
        return self._neighbour

    @property
    def diag(self) -> torch.Tensor:
        return self._diag

    @diag.setter
    def diag(self, value: torch.Tensor) -> None:
        self._diag = value.to(device=self._device, dtype=self._dtype)
        self.invalidate_cache()

    @property
    def lower(self) -> torch.Tensor:
        return self._lower

    @lower.setter
    def lower(self, value: torch.Tensor) -> None:
        self._lower = value.to(device=self._device, dtype=self._dtype)
        self.invalidate_cache()

    @property
    def upper(self) -> torch.Tensor:
        return self._upper

    @upper.setter
    def upper(self, value: torch.Tensor) -> None:
        self._upper = value.to(device=self._device, dtype=self._dtype)
        self.invalidate_cache()

    # ------------------------------------------------------------ native state
    @property
    def addressing(self) -> LduAddressing:
        """Native addressing handle (built on first use, shared per topology)."""
        if self._addr is None:
            require_gpu(self._owner)
            self._addr = addressing_for(self._n_cells, self._owner_src, self._neighbour_src,
                                        self._device)
        return self._addr

    def invalidate_cache(self) -> None:
        self._packed.clear()
        self._csr_cache.clear()

    def _check_compute(self) -> None:
        require_gpu(self._diag)
        if self._dtype != torch.float64:
            raise TypeError("libpfm LDU kernels compute in float64; use mixed-precision solvers "
                            "for float32 inner work")

    def coefficients(self) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """(diag, lower, upper), detached and contiguous.

        The caller binds the tuple for as long as the device pointers are in
        use: a contiguous copy of a strided coefficient is a new tensor, and
        one freed before its kernel runs can be handed by the caching
        allocator to the next copy, so two arguments would share one block.
        """
        return (self._diag.detach().contiguous(), self._lower.detach().contiguous(),
                self._upper.detach().contiguous())

    def packed(self, transpose: bool = False) -> torch.Tensor:
        """Row-slice packed off-diagonal coefficients (cached)."""
        self._check_compute()
        lo, up = self._lower.detach(), self._upper.detach()
        key = (bool(transpose), lo.data_ptr(), up.data_ptr(), lo._version, up._version)
        hit = self._packed.get(bool(transpose))
        if hit is not None and hit[0] == key:
            return hit[1]
        addr = self.addressing
        buf = torch.empty(max(addr.n_slots, 1), device=self._device, dtype=torch.float64)
        loc, upc = lo.contiguous(), up.contiguous()
        _lib.check(_lib.lib().pfm_ldu_pack(addr.handle, int(transpose), _lib.ptr(loc),
                                           _lib.ptr(upc), _lib.ptr(buf),
                                           current_stream(self._device)), "pfm_ldu_pack")
        self._packed[bool(transpose)] = (key, buf)
        return buf

    def packed32(self, transpose: bool = False) -> torch.Tensor:
        """FP32 copy of the packed coefficients (mixed-precision solves, cached)."""
        self._check_compute()
        lo, up = self._lower.detach(), self._upper.detach()
        key = ("f32", bool(transpose), lo.data_ptr(), up.data_ptr(), lo._version, up._version)
        hit = self._packed.get(("f32", bool(transpose)))
        if hit is not None and hit[0] == key:
            return hit[1]
        addr = self.addressing
        buf = torch.empty(max(addr.n_slots, 1), device=self._device, dtype=torch.float32)
        loc, upc = lo.contiguous(), up.contiguous()
        _lib.check(_lib.lib().pfm_ldu_pack_f32(addr.handle, int(transpose),
                                               _lib.ptr(loc), _lib.ptr(upc),
                                               _lib.ptr(buf), current_stream(self._device)),
                   "pfm_ldu_pack_f32")
        self._packed[("f32", bool(transpose))] = (key, buf)
        return buf

    # ------------------------------------------------------------ products
    def _check_x(self, x: torch.Tensor) -> torch.Tensor:
        assert_floating(x, "x")
        if x.shape[0] != self._n_cells:
            raise ValueError(f"x has {x.shape[0]} elements, expected {self._n_cells}")
        return x.to(device=self._device, dtype=self._dtype).contiguous()

    def _product(self, x: torch.Tensor, transpose: int) -> torch.Tensor:
        x = self._check_x(x)
        self._check_compute()
        y = torch.empty(self._n_cells, device=self._device, dtype=self._dtype)
        if self._n_cells == 0:
            return y
        d = self._diag.detach().contiguous()
        lo = self._lower.detach().contiguous()
        up = self._upper.detach().contiguous()
        _lib.check(_lib.lib().pfm_spmv_ldu(self.addressing.handle, transpose, _lib.ptr(d),
                                           _lib.ptr(lo), _lib.ptr(up), _lib.ptr(x.detach()),
                                           _lib.ptr(y), current_stream(self._device)),
                   "pfm_spmv_ldu")
        return y

    def Ax(self, x: torch.Tensor) -> torch.Tensor:
        """y = A x (reference ``ldu_matrix.py:172-217``)."""
        return self._product(x, 0)

    def Ax_T(self, x: torch.Tensor) -> torch.Tensor:
        """y = A^T x (reference ``ldu_matrix.py:219-258``)."""
        return self._product(x, 1)

    def Ax_packed(self, x: torch.Tensor, transpose: bool = False) -> torch.Tensor:
        """y = A x (or the transpose matrix) through the cached packed copy."""
        x = self._check_x(x)
        pk = self.packed(transpose)
        y = torch.empty(self._n_cells, device=self._device, dtype=self._dtype)
        if self._n_cells:
            d = self._diag.detach().contiguous()
            _lib.check(_lib.lib().pfm_spmv(self.addressing.handle, _lib.ptr(d),
                                           _lib.ptr(pk), _lib.ptr(x.detach()), _lib.ptr(y),
                                           current_stream(self._device)), "pfm_spmv")
        return y

    # ------------------------------------------------------------ sparse forms
    def to_sparse_coo(self) -> torch.Tensor:
        """Coalesced COO tensor of A (``ldu_matrix.py:278-305``): entries
        (i, i) = diag, (owner, neighbour) = lower, (neighbour, owner) = upper."""
        n = self._n_cells
        ar = torch.arange(n, device=self._device, dtype=torch.int64)
        own, nei = self._owner.long(), self._neighbour.long()
        if self._n_internal_faces:
            idx = torch.cat([torch.stack([ar, ar]), torch.stack([own, nei]),
                             torch.stack([nei, own])], dim=1)
            val = torch.cat([self._diag, self._lower, self._upper])
        else:
            idx, val = torch.stack([ar, ar]), self._diag
        return torch.sparse_coo_tensor(idx, val, (n, n), device=self._device).coalesce()

    def to_sparse_csr(self) -> torch.Tensor:
        """CSR tensor of A (``ldu_matrix.py:307-313``)."""
        return self.to_sparse_coo().to_sparse_csr()

    def to_sparse_csr_cached(self) -> torch.Tensor:
        """CSR tensor of A, cached until a coefficient setter or
        ``invalidate_cache`` (``ldu_matrix.py:315-333``)."""
        hit = self._csr_cache.get("csr")
        if hit is not None and hit.device == self._device:
            return hit
        csr = self.to_sparse_csr()
        self._csr_cache["csr"] = csr
        return csr

    def Ax_sparse(self, x: torch.Tensor) -> torch.Tensor:
        """y = A x for x of shape (n,) or (n, k) (``ldu_matrix.py:345-364``)."""
        assert_floating(x, "x")
        if x.dim() not in (1, 2) or x.shape[0] != self._n_cells:
            raise ValueError(f"x has shape {tuple(x.shape)}, expected ({self._n_cells},) or "
                             f"({self._n_cells}, k)")
        x = x.to(device=self._device, dtype=self._dtype)
        if x.dim() == 1:
            return self.Ax_packed(x)
        self._check_compute()
        k = int(x.shape[1])
        xc = x.detach().contiguous()
        y = torch.empty(self._n_cells, k, device=self._device, dtype=self._dtype)
        if self._n_cells and k:
            d = self._diag.detach().contiguous()
            pk = self.packed(False)
            _lib.check(_lib.lib().pfm_spmv_multi(self.addressing.handle, _lib.ptr(d), _lib.ptr(pk),
                                                 k, _lib.ptr(xc), _lib.ptr(y),
                                                 current_stream(self._device)), "pfm_spmv_multi")
        return y

    def Ax_batched(self, x: torch.Tensor) -> torch.Tensor:
        """Y = A X for an (n, k) block of right-hand sides (``ldu_matrix.py:366-381``)."""
        return self.Ax_sparse(x)

    def __repr__(self) -> str:
        return (f"LduMatrix(n_cells={self._n_cells}, n_internal_faces={self._n_internal_faces}, "
                f"device={self._device}, dtype={self._dtype})")

    def add_to_diag(self, values: torch.Tensor) -> None:
        if values.dim() == 0:
            self._diag = self._diag + values
        else:
            self._diag = self._diag + values.to(device=self._device, dtype=self._dtype)
        self.invalidate_cache()

    def transposed(self) -> "LduMatrix":
        """A^T sharing this topology (lower and upper swapped)."""
        m = LduMatrix(self._n_cells, self._owner_src, self._neighbour_src, device=self._device,
                      dtype=self._dtype)
        m._addr = self._addr
        m._diag, m._lower, m._upper = self._diag, self._upper, self._lower
        return m

    def __repr__(self) -> str:
        return (f"LduMatrix(n_cells={self._n_cells}, n_internal_faces={self._n_internal_faces}, "
                f"device={self._device}, dtype={self._dtype})")
