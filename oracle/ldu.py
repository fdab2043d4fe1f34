"""LDU products and Laplacian assembly (oracle; test infrastructure only).

``ax``/``ax_t`` restate ``core/ldu_matrix.py:172-258``: y = diag*x, then the
owner-side scatter, then the neighbour-side scatter, each scatter summing its
contributions per row in face order from 0 (``core/backend.py:34-64``
``Tensor.scatter_add`` on CPU).  ``np.bincount`` with weights accumulates in
input order from 0, so the association and rounding are the reference's.

``laplacian`` restates ``fvm.laplacian`` (``discretisation/operators.py:490-517``)
and ``dirichlet_diag`` the laplacianFoam patch term
(``applications/laplacian_foam.py:342-366``).
"""

from __future__ import annotations

import numpy as np


def _scatter(vals: np.ndarray, idx: np.ndarray, n: int) -> np.ndarray:
    return np.bincount(idx, weights=vals, minlength=n).astype(np.float64)


def ax(diag, lower, upper, owner, neighbour, x) -> np.ndarray:
    n = diag.shape[0]
    y = diag * x
    if len(owner):
        x_own = x[owner]
        x_nei = x[neighbour]
        y = y + _scatter(lower * x_nei, owner, n)
        y = y + _scatter(upper * x_own, neighbour, n)
    return y


def ax_t(diag, lower, upper, owner, neighbour, x) -> np.ndarray:
    n = diag.shape[0]
    y = diag * x
    if len(owner):
        x_own = x[owner]
        x_nei = x[neighbour]
        y = y + _scatter(lower * x_own, neighbour, n)
        y = y + _scatter(upper * x_nei, owner, n)
    return y


def dense(diag, lower, upper, owner, neighbour) -> np.ndarray:
    """Dense A with A[own,nei] = lower, A[nei,own] = upper (ldu_matrix.py:65-71)."""
    n = diag.shape[0]
    a = np.diag(np.asarray(diag, dtype=np.float64)).copy()
    np.add.at(a, (owner, neighbour), lower)
    np.add.at(a, (neighbour, owner), upper)
    return a


def face_mag(face_areas) -> np.ndarray:
    """|S_f| as the reference computes it (torch norm over dim 1, FMA order)."""
    from ._clib import clib, f64, p
    s = f64(face_areas)
    out = np.empty(s.shape[0])
    clib().orc_norm3(s.shape[0], p(s), p(out))
    return out


def laplacian(gamma_int, mag_sf_int, delta_int, vol, owner, neighbour):
    """Internal-face part of fvm.laplacian: returns (diag, lower, upper).

    face_coeff = D*|S|*delta (left to right); lower = -face_coeff/V_P;
    upper = -face_coeff/V_N; diag = 0 + scatter(c/V_P, own) + scatter(c/V_N, nei).
    The generic boundary term of fvm.laplacian adds D*|S|*0/V = 0 because the
    reference's boundary delta is 0 (mesh_geometry.py:376-393).
    """
    n = vol.shape[0]
    c = gamma_int * mag_sf_int * delta_int
    vp = vol[owner]
    vn = vol[neighbour]
    lower = -c / vp
    upper = -c / vn
    diag = np.zeros(n)
    diag = diag + _scatter(c / vp, owner, n)
    diag = diag + _scatter(c / vn, neighbour, n)
    return diag, lower, upper


def dirichlet_diag(diag, gamma_b, mag_b, delta_b, vol, cells):
    """laplacianFoam fixedValue term: diag + scatter(D*|S_b|*delta_b / V_P)."""
    n = diag.shape[0]
    coeff = gamma_b * mag_b * delta_b
    return diag + _scatter(coeff / vol[cells], cells, n)


def dirichlet_diag_patches(diag, patches, vol):
    """The patch loop of ``applications/laplacian_foam.py:318-366``: each patch
    (gamma_b, |S_b|, delta_b, cells), faces in face order, adds its own
    scatter_add to the diagonal in turn (:357)."""
    for gamma_b, mag_b, delta_b, cells in patches:
        diag = dirichlet_diag(diag, gamma_b, mag_b, delta_b, vol, cells)
    return diag
