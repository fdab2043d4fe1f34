"""pfm_pcg_solve_noalloc (the PCG entry point that allocates nothing) against
pfm_pcg_solve with a caller workspace, through the C ABI: the same iteration
count, status, residual history and solution bits, on two calls in a row, the
second on a smaller system (the shared device array keeps the first call's
contents); a request beyond the array's size is rejected before anything is
enqueued."""

import ctypes as C

import numpy as np
import pytest
import torch

from pyopenfoam_amd import LduMatrix, _lib
from pyopenfoam_amd.core.device import current_stream
from pyopenfoam_amd.solvers import DICPreconditioner
from tests.helpers import bits_equal, system

pytestmark = pytest.mark.gpu


def _setup(n, dev):
    c, d, lo, up = system(n, 0.2, 1.0)
    m = LduMatrix(c.n_cells, torch.from_numpy(c.owner), torch.from_numpy(c.neighbour), device=dev)
    m.diag, m.lower, m.upper = (torch.from_numpy(v) for v in (d, lo, up))
    rD = DICPreconditioner(m).diagonal
    return c, m, rD


def _solve(m, rD, b, noalloc, max_iter=5000):
    L = _lib.lib()
    dev = m.device
    prm = _lib.SolveParams(1e-300, 1e-8, max_iter, 0, _lib.PFM_PRECOND_DIAG, 0)
    res = _lib.SolveResult()
    hist = np.zeros(max_iter + 1)
    x = torch.zeros(m.n_cells, dtype=torch.float64, device=dev)
    d, pk = m.diag.detach().contiguous(), m.packed(False)
    h = m.addressing.handle
    if noalloc:
        rc = L.pfm_pcg_solve_noalloc(h, C.byref(prm), _lib.ptr(d), _lib.ptr(pk), _lib.ptr(rD),
                                     _lib.ptr(b), _lib.ptr(x), hist.ctypes.data, C.byref(res),
                                     current_stream(dev))
    else:
        work = torch.empty(int(L.pfm_pcg_workspace_bytes(h, max_iter)), dtype=torch.uint8,
                           device=dev)
        rc = L.pfm_pcg_solve(h, C.byref(prm), _lib.ptr(d), _lib.ptr(pk), _lib.ptr(rD), _lib.ptr(b),
                             _lib.ptr(x), _lib.ptr(work), hist.ctypes.data, C.byref(res),
                             current_stream(dev))
    _lib.check(rc, "pcg")
    return x.cpu().numpy(), res.iterations, res.status, hist[:res.iterations].copy()


def test_noalloc_pcg_matches_workspace_pcg(gpu_device):
    assert _lib.lib().pfm_pcg_noalloc_max_bytes() == 68 << 20
    for n in (32, 12):  # the second call on a smaller system
        print(f"noalloc pcg n={n}", flush=True)
        c, m, rD = _setup(n, gpu_device)
        b = torch.from_numpy(c.b).to(gpu_device)
        x0, it0, st0, h0 = _solve(m, rD, b, noalloc=False)
        x1, it1, st1, h1 = _solve(m, rD, b, noalloc=True)
        assert st0 == st1 == _lib.PFM_SOLVE_CONVERGED
        assert it0 == it1, (it0, it1)
        assert bits_equal(h0, h1)
        assert bits_equal(x0, x1)


def test_noalloc_pcg_rejects_oversized_requests(gpu_device):
    c, m, rD = _setup(8, gpu_device)
    L = _lib.lib()
    cap = int(L.pfm_pcg_noalloc_max_bytes())
    big = (cap // 16) + 1  # the residual history alone exceeds the array
    assert int(L.pfm_pcg_workspace_bytes(m.addressing.handle, big)) > cap
    b = torch.from_numpy(c.b).to(gpu_device)
    prm = _lib.SolveParams(1e-300, 1e-8, big, 0, _lib.PFM_PRECOND_DIAG, 0)
    res = _lib.SolveResult()
    x = torch.zeros(m.n_cells, dtype=torch.float64, device=gpu_device)
    rc = L.pfm_pcg_solve_noalloc(m.addressing.handle, C.byref(prm),
                                 _lib.ptr(m.diag.detach().contiguous()), _lib.ptr(m.packed(False)),
                                 _lib.ptr(rD), _lib.ptr(b), _lib.ptr(x), None, C.byref(res),
                                 current_stream(gpu_device))
    assert rc == _lib.PFM_ERR_ARG
    assert b"no-allocation" in L.pfm_last_error() or b"holds at most" in L.pfm_last_error()
    assert not x.any()
