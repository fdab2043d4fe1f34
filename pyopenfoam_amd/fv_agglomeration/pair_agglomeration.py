"""PairGamgAgglomeration (reference ``fv_agglomeration/pair_agglomeration.py``).

``PairGamgAgglomeration(n_cells, owner, neighbour)``, ``build_levels(n_levels=5)``
(:47-88), ``n_levels`` and ``levels`` (:133-141).  Each level is the
reference's dict {level, n_cells_fine, n_cells_coarse, cell_map}, cell_map
mapping the finest cells to that level (cumulative, int64).  The pairing of
``_pair_cells`` (:90-131: faces in order, merge the two current aggregates
when both are unpaired; renumber by sorted representative; stop when no pair
forms, at n <= 4 or when a level keeps more than 90 % of the cells) runs in
the native ``pfm_pair_agglomerate`` (host C++, O(faces) per level instead of
the reference's O(cells * pairs)); the maps are the same.
"""

from __future__ import annotations

import numpy as np
import torch

from .. import _lib

__all__ = ["PairGamgAgglomeration"]


class PairGamgAgglomeration:
    def __init__(self, n_cells: int, owner: torch.Tensor, neighbour: torch.Tensor):
        self._n_cells = int(n_cells)
        self._owner = owner
        self._neighbour = neighbour
        self._levels: list[dict] = []

    def build_levels(self, n_levels: int = 5) -> list[dict]:
        import ctypes as C
        n = self._n_cells
        nfi = int(self._neighbour.shape[0])
        own = _lib.host_i32(self._owner)[:nfi]
        nei = _lib.host_i32(self._neighbour)
        k = max(int(n_levels), 0)
        maps = np.empty((max(k, 1), max(n, 1)), dtype=np.int32)
        counts = np.empty(max(k, 1), dtype=np.int32)
        made = C.c_int32(0)
        if k and n:
            _lib.check(_lib.lib().pfm_pair_agglomerate(n, nfi, _lib.ptr(own), _lib.ptr(nei), k,
                                                       _lib.ptr(maps), _lib.ptr(counts),
                                                       C.byref(made)), "pfm_pair_agglomerate")
        levels, fine = [], n
        for lev in range(made.value):
            levels.append({"level": lev, "n_cells_fine": fine,
                           "n_cells_coarse": int(counts[lev]),
                           "cell_map": torch.from_numpy(maps[lev, :n].astype(np.int64))})
            fine = int(counts[lev])
        self._levels = levels
        return levels

    @property
    def n_levels(self) -> int:
        return len(self._levels)

    @property
    def levels(self) -> list[dict]:
        return self._levels
