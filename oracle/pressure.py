"""Pressure-equation step (oracle; test infrastructure only).

Restates, in numpy and in the reference's operation order:

* ``face_weights``      -- mesh/mesh_geometry.py:262-327 (compute_face_weights)
* ``assemble``          -- solvers/pressure_equation.py:49-152
                           (assemble_pressure_equation): (diag, lower, upper, source)
* ``correct_face_flux`` -- pressure_equation.py:270-332
* ``correct_velocity``  -- pressure_equation.py:187-267
* ``adjust_phi``        -- pressure_equation.py:335-377

``np.add.at`` accumulates repeated indices one after another in index order,
as torch's CPU scatter_add / index_add_ do; |v| of a 3-vector uses the FMA
order of torch's norm (``oracle.ldu.face_mag``).  Pins (tests/test_oracle_golden.py):
A_p = 1 reduces the matrix to the golden-pinned Laplacian bit for bit; the
divergence of a uniform flux field S.U0 vanishes; a unit flux on one face
lands on its two cells with opposite signs; a linear pressure gives the exact
Gauss gradient in interior cells of a uniform mesh and the exact flux
correction -(1/A_p) g.S_f there; closed-domain adjustment zeroes the net
boundary flux.
"""

from __future__ import annotations

import numpy as np

from .ldu import face_mag


def _inv(a_p):
    a = np.abs(np.asarray(a_p, dtype=np.float64))
    return 1.0 / np.where(a < 1e-30, 1e-30, a)


def face_weights(face_centres, cell_centres, owner, neighbour):
    fc = np.asarray(face_centres, dtype=np.float64)
    nf, nfi = fc.shape[0], len(neighbour)
    w = np.ones(nf)
    if nfi:
        own = np.asarray(owner)[:nfi]
        dp = face_mag(fc[:nfi] - cell_centres[own])
        dn = face_mag(fc[:nfi] - cell_centres[np.asarray(neighbour)])
        den = dp + dn
        w[:nfi] = dn / np.where(den > 1e-30, den, 1.0)
    return w


def _inv_face(a_p, w_int, own, nei):
    inv = _inv(a_p)
    return w_int * inv[own] + (1.0 - w_int) * inv[nei]


def assemble(phi, a_p, w, mag_int, delta_int, vol, owner, neighbour):
    """phi (n_faces,), w/mag/delta over internal faces -> (diag, lower, upper, source)."""
    nfi = len(neighbour)
    n = vol.shape[0]
    own = np.asarray(owner)[:nfi]
    nei = np.asarray(neighbour)
    c = (_inv_face(a_p, w[:nfi], own, nei) * mag_int) * delta_int
    lower = (-c) / vol[own]
    upper = (-c) / vol[nei]
    s1 = np.zeros(n)
    np.add.at(s1, own, c / vol[own])
    s2 = np.zeros(n)
    np.add.at(s2, nei, c / vol[nei])
    diag = (np.zeros(n) + s1) + s2
    t1 = np.zeros(n)
    np.add.at(t1, own, -phi[:nfi])
    t2 = np.zeros(n)
    np.add.at(t2, nei, phi[:nfi])
    src = (np.zeros(n) + t1) + t2
    if phi.shape[0] > nfi:
        t3 = np.zeros(n)
        np.add.at(t3, np.asarray(owner)[nfi:], -phi[nfi:])
        src = src + t3
    return diag, lower, upper, src / vol


def correct_face_flux(phi, p, a_p, w, mag_int, delta_int, owner, neighbour):
    nfi = len(neighbour)
    own = np.asarray(owner)[:nfi]
    nei = np.asarray(neighbour)
    dp = p[own] - p[nei]
    out = np.array(phi, dtype=np.float64, copy=True)
    out[:nfi] = phi[:nfi] + ((_inv_face(a_p, w[:nfi], own, nei) * dp) * mag_int) * delta_int
    return out


def correct_velocity(hbya, p, a_p, w, face_areas, vol, owner, neighbour, alpha=1.0):
    nfi = len(neighbour)
    n = vol.shape[0]
    owner = np.asarray(owner)
    own = owner[:nfi]
    nei = np.asarray(neighbour)
    pf = w[:nfi] * p[own] + (1.0 - w[:nfi]) * p[nei]
    contrib = pf[:, None] * face_areas[:nfi]
    g = np.zeros((n, 3))
    np.add.at(g, own, contrib)
    np.add.at(g, nei, -contrib)
    if face_areas.shape[0] > nfi:
        bo = owner[nfi:]
        np.add.at(g, bo, p[bo][:, None] * face_areas[nfi:])
    v = np.where(vol < 1e-30, 1e-30, vol)
    g = g / v[:, None]
    u = hbya - (alpha * _inv(a_p))[:, None] * g
    m = face_mag(u)
    m = np.where(m < 1e-30, 1e-30, m)
    return np.where((m > 1e4)[:, None], u * (1e4 / m)[:, None], u)


def adjust_phi(phi, n_internal, closed=True):
    """Returns (adjusted, phi_out); the net flux is summed with math.fsum
    (exactly rounded), one of the orders torch's sum may take."""
    import math
    phi = np.array(phi, dtype=np.float64, copy=True)
    nb = phi.shape[0] - n_internal
    if nb <= 0 or not closed:
        return False, phi
    net = math.fsum(phi[n_internal:].tolist())
    if abs(net) < 1e-30:
        return False, phi
    phi[n_internal:] = phi[n_internal:] - net / nb
    return True, phi
