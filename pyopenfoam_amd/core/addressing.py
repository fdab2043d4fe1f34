"""Native LDU addressing handles (``pfm_ldu``) and their cache.

A handle holds the device-side addressing of one topology: the row-slice
layout used by every SpMV, the DIC cell levels and the DILU face levels.  It
is built once per (n_cells, owner, neighbour) and shared by every matrix with
that topology (the forward and adjoint matrices of a differentiable solve,
GAMG's fine level, repeated solves in a time loop).
"""

from __future__ import annotations

import ctypes as C
import threading
from collections import OrderedDict

import numpy as np
import torch

from .. import _lib

__all__ = ["LduAddressing", "DirichletPlan", "addressing_for"]


class LduAddressing:
    def __init__(self, n_cells: int, owner, neighbour, device: torch.device):
        own = _lib.host_i32(owner)
        nei = _lib.host_i32(neighbour)
        if own.shape != nei.shape:
            raise ValueError("owner and neighbour must have the same length")
        self.n_cells = int(n_cells)
        self.n_faces = int(own.shape[0])
        self.device = torch.device(device)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().pfm_ldu_create(self.n_cells, self.n_faces, _lib.ptr(own),
                                                 _lib.ptr(nei), C.byref(h)), "pfm_ldu_create")
        self._h = h
        info = _lib.LduInfo()
        _lib.check(_lib.lib().pfm_ldu_get_info(h, C.byref(info)))
        self.info = info
        self.n_slots = int(info.n_slots)
        self.owner_host = own
        self.neighbour_host = nei
        # int32 device copies for callers that need face addressing
        self.owner32 = torch.from_numpy(own).to(self.device)
        self.neighbour32 = torch.from_numpy(nei).to(self.device)

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def set_progress(self, fn=None, interval_s: float = 20.0) -> None:
        """Report long device solves on this topology: ``fn(iterations)`` is
        called from the solving thread at most every ``interval_s`` seconds.
        ``fn=None`` turns reports off."""
        if fn is None:
            self._progress = None
            _lib.check(_lib.lib().pfm_ldu_set_progress(self._h, None, None, 0.0))
            return
        self._progress = _lib.PROGRESS_FN(lambda it, user: fn(int(it)))  # kept alive here
        _lib.check(_lib.lib().pfm_ldu_set_progress(self._h, self._progress, None,
                                                   float(interval_s)))

    def dirichlet_plan(self, cells, patch_sizes=None) -> "DirichletPlan":
        """Plan of the Dirichlet faces for ``pfm_laplacian_assemble``.

        ``cells``: the owner cell of each Dirichlet face (numpy or torch), in
        the reference's order -- patch after patch, faces in face order;
        ``patch_sizes``: faces per patch (default: one patch).  Plans are
        cached on this handle by the cells' values and the patch sizes.
        """
        c = _lib.host_i32(cells)
        sizes = (int(c.shape[0]),) if patch_sizes is None else tuple(int(k) for k in patch_sizes)
        if sum(sizes) != c.shape[0]:
            raise ValueError("patch sizes do not add up to the number of Dirichlet faces")
        key = (sizes, c.tobytes())
        plans = self.__dict__.setdefault("_dir_plans", OrderedDict())
        with _cache_lock:
            hit = plans.get(key)
            if hit is not None:
                plans.move_to_end(key)
                return hit
        plan = DirichletPlan(self, c, sizes)
        with _cache_lock:
            plans[key] = plan
            while len(plans) > _CACHE_SIZE:
                plans.popitem(last=False)
        return plan

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pfm_ldu_destroy(h)
            except Exception:
                pass
            self._h = None


class DirichletPlan:
    """``pfm_dirichlet`` handle: Dirichlet faces grouped by cell in the library."""

    def __init__(self, addr: LduAddressing, cells: np.ndarray, patch_sizes: tuple):
        self.addr = addr  # the topology handle outlives the plan
        self.n_dir = int(cells.shape[0])
        self.patch_sizes = patch_sizes
        ptr = np.zeros(len(patch_sizes) + 1, dtype=np.int32)
        ptr[1:] = np.cumsum(patch_sizes)
        h = C.c_void_p()
        with torch.cuda.device(addr.device):
            _lib.check(_lib.lib().pfm_dirichlet_create(addr.handle, self.n_dir, _lib.ptr(cells),
                                                       len(patch_sizes), _lib.ptr(ptr),
                                                       C.byref(h)), "pfm_dirichlet_create")
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pfm_dirichlet_destroy(h)
            except Exception:
                pass
            self._h = None


_cache: "OrderedDict[tuple, tuple]" = OrderedDict()
_cache_lock = threading.Lock()
_CACHE_SIZE = 8


def addressing_for(n_cells: int, owner: torch.Tensor, neighbour: torch.Tensor,
                   device: torch.device) -> LduAddressing:
    """Cached handle for a topology.

    Keyed by the identity and version counter of the caller's owner/neighbour
    tensors; the cache keeps those tensors alive so the key cannot be reused
    by a different topology.
    """
    key = (int(n_cells), id(owner), id(neighbour), owner._version, neighbour._version,
           str(device))
    with _cache_lock:
        hit = _cache.get(key)
        if hit is not None and hit[1] is owner and hit[2] is neighbour:
            _cache.move_to_end(key)
            return hit[0]
    addr = LduAddressing(n_cells, owner, neighbour, device)
    with _cache_lock:
        _cache[key] = (addr, owner, neighbour)
        while len(_cache) > _CACHE_SIZE:
            _cache.popitem(last=False)
    return addr


def clear_cache() -> None:
    with _cache_lock:
        _cache.clear()
