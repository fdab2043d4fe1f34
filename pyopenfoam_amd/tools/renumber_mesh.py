"""renumberMesh: Reverse Cuthill-McKee cell order (reference
``tools/renumber_mesh.py``).

``renumber_mesh(mesh) -> RenumberResult`` (:62-99) with the reference's
fields: ``permutation`` (perm[new] = old), ``inverse_permutation``
(inv[old] = new), ``original_bandwidth`` and ``renumbered_bandwidth``.  The
order comes from the native ``pfm_rcm`` (host C++, O(cells + faces)) with the
tie-breaks of ``_reverse_cuthill_mckee`` (:163-239).  The bandwidths are those
of ``_bandwidth`` (:258-275) on the symmetric adjacency and on
``_apply_permutation_to_adj`` (:278-305), which maps each edge (i, j) to
(perm[i], perm[j]); that reproduces the reference's renumbered figure
exactly (it indexes ``perm`` with the old cell ids).  The mesh is not
modified.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib

__all__ = ["RenumberResult", "renumber_mesh"]


@dataclass
class RenumberResult:
    permutation: torch.Tensor
    inverse_permutation: torch.Tensor
    original_bandwidth: int
    renumbered_bandwidth: int


def _host_index(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.int64)


def renumber_mesh(mesh) -> RenumberResult:
    """RCM permutation of ``mesh`` (anything with n_cells, n_internal_faces,
    owner and neighbour, as the reference's FvMesh)."""
    n = int(mesh.n_cells)
    nfi = int(mesh.n_internal_faces)
    own = _host_index(mesh.owner)[:nfi]
    nei = _host_index(mesh.neighbour)[:nfi]
    o32, q32 = _lib.host_i32(own), _lib.host_i32(nei)
    perm = np.empty(n, dtype=np.int32)
    _lib.check(_lib.lib().pfm_rcm(n, nfi, _lib.ptr(o32), _lib.ptr(q32), _lib.ptr(perm)),
               "pfm_rcm")
    perm64 = perm.astype(np.int64)
    inv = np.empty(n, dtype=np.int64)
    inv[perm64] = np.arange(n, dtype=np.int64)
    bw0 = int(np.abs(own - nei).max()) if nfi else 0
    bw1 = int(np.abs(perm64[own] - perm64[nei]).max()) if nfi else 0
    return RenumberResult(permutation=torch.from_numpy(perm64),
                          inverse_permutation=torch.from_numpy(inv),
                          original_bandwidth=bw0, renumbered_bandwidth=bw1)
