"""Adjoint of the linear solve (oracle; test infrastructure only).

``ldu_grads`` restates DifferentiableLinearSolve.backward
(differentiable/linear_solver.py:99-151): the transposed matrix is the LDU
with lower and upper swapped, lambda solves it with the forward solver from a
zero initial guess, and
  grad_diag = -lam*x, grad_lower = -lam[own]*x[nei], grad_upper = -lam[nei]*x[own],
  grad_b = lam.
``gamma_grad`` chains those through the assembly of fvm.laplacian
(discretisation/operators.py:490-517) and the laplacianFoam Dirichlet term,
the derivative torch autograd takes through those ops:
  dJ/dgamma_f = |S_f| delta_f ((g_diag[P] - g_lower_f)/V_P + (g_diag[N] - g_upper_f)/V_N)
  dJ/dgamma_b = |S_b| delta_b g_diag[P] / V_P.
"""

from __future__ import annotations

import numpy as np


def ldu_grads(lam, x, owner, neighbour):
    return -lam * x, -lam[owner] * x[neighbour], -lam[neighbour] * x[owner]


def gamma_grad(g_diag, g_lower, g_upper, mag_sf, delta, vol, owner, neighbour,
               dir_cells=None, dir_mag=None, dir_delta=None):
    sd = mag_sf * delta
    g = sd * ((g_diag[owner] - g_lower) / vol[owner] + (g_diag[neighbour] - g_upper) / vol[neighbour])
    gb = None
    if dir_cells is not None:
        gb = dir_mag * dir_delta * g_diag[dir_cells] / vol[dir_cells]
    return g, gb
