"""GAMG solver on the GPU (reference ``solvers/gamg.py``).

``GAMGSolver(tolerance, rel_tol, max_iter=100, min_iter, verbose, n_pre_smooth=2,
n_post_smooth=2, max_levels=10, min_cells_coarse=10, smoother="PCG")`` has the
reference's semantics (:133-585): greedy aggregation with target max(n//2,
min_cells_coarse), the non-Galerkin coarse matrix of ``_build_coarse_matrix``
(aggregate-summed diagonal, first-seen coarse faces, couplings inside an
aggregate dropped), V-cycles of undamped Jacobi (1/|diag|) with 50 sweeps on
the coarsest level, residual monitored after every cycle with no initial
residual.  The ``smoother`` argument is stored and ignored, as in the
reference (:157).

Additive keywords (not in the reference): ``agglomeration="pair"`` builds the
levels with PairGamgAgglomeration's pairing
(``fv_agglomeration/pair_agglomeration.py``) and ``coarse="galerkin"`` adds the
intra-aggregate couplings to the coarse diagonal (the Galerkin product of the
piecewise-constant prolongation), which the reference drops;
``level_smoother="GaussSeidel"`` replaces every Jacobi smoothing step (pre,
post and coarsest) by a symmetric Gauss-Seidel sweep with the arithmetic of
``GaussSeidelSmoother`` (``solvers/smoothers.py:92-149``);
``precision="mixed"`` solves in FP64 iterative refinement whose inner V-cycles
run in FP32 on the finest ``fp32_levels`` levels (default: every level
above 4096 cells; FP64 on the coarser ones, the coarsest-level solve
included) to ``inner_rel_tol`` (default 0.1: the FP32 residual of a
stationary cycle stalls at its rounding floor, which a tighter inner
tolerance would not see past; on a 256^3 lognormal(0, 1.5) system 1e-2
already lies near that floor and takes 1.67x the V-cycles, while on
128^3 lognormal(0, 1) 0.1 costs 5 % more time than 1e-2); the residual
monitor then runs on the FP64 true residual after each refinement step
(the initial residual included), ``max_iter`` bounds the total V-cycles
and ``last_info.inner_iterations`` counts them.

The hierarchy topology is built once per mesh topology (host C++) and cached;
coefficients are coarsened on the device at every solve, and the whole V-cycle
runs in libpfm kernels (``pfm_gamg_solve``).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from ..core.device import current_stream, require_gpu
from ..core.ldu_matrix import LduMatrix
from ._krylov import KrylovWorkspace
from .linear_solver import LinearSolverBase
from .residual import ConvergenceInfo, ResidualMonitor

__all__ = ["GAMGSolver", "GamgHierarchy"]

_AGGLOM = {"greedy": _lib.PFM_AGGLOM_GREEDY, "faceareapair": _lib.PFM_AGGLOM_PAIR,
           "pair": _lib.PFM_AGGLOM_PAIR}
_COARSE = {"reference": _lib.PFM_COARSE_REFERENCE, "galerkin": _lib.PFM_COARSE_GALERKIN}
_LEVEL_SMOOTHER = {"jacobi": _lib.PFM_SMOOTHER_JACOBI, "gaussseidel": _lib.PFM_SMOOTHER_GAUSS_SEIDEL,
                   "gs": _lib.PFM_SMOOTHER_GAUSS_SEIDEL}


class GamgHierarchy:
    """Native hierarchy handle (``pfm_gamg``) for one topology and parameter set."""

    def __init__(self, addressing, params: _lib.GamgParams):
        h = C.c_void_p()
        _lib.check(_lib.lib().pfm_gamg_create(addressing.handle, C.byref(params), C.byref(h)),
                   "pfm_gamg_create")
        self._h = h
        self._addr = addressing  # the fine handle must outlive the hierarchy
        self.n_levels = _lib.lib().pfm_gamg_num_levels(h)

    @property
    def handle(self):
        return self._h

    def level_sizes(self) -> list[tuple[int, int]]:
        out = []
        for k in range(self.n_levels):
            n, f = C.c_int32(), C.c_int32()
            _lib.check(_lib.lib().pfm_gamg_level_info(self._h, k, C.byref(n), C.byref(f)))
            out.append((n.value, f.value))
        return out

    def level_maps(self, level: int):
        sizes = self.level_sizes()
        n, nf = sizes[level]
        nc, nfc = sizes[level + 1]
        f2c = np.empty(n, np.int32)
        fmap = np.empty(nf, np.int32)
        co = np.empty(nfc, np.int32)
        cn = np.empty(nfc, np.int32)
        _lib.check(_lib.lib().pfm_gamg_level_maps(self._h, level, f2c.ctypes.data, fmap.ctypes.data,
                                                  co.ctypes.data, cn.ctypes.data))
        return f2c, fmap, co, cn

    def level_topology(self, level: int) -> tuple[np.ndarray, np.ndarray]:
        """(owner, neighbour) of a level: the fine handle's on level 0, the
        coarse faces built from the previous level's maps above it."""
        if level == 0:
            return self._addr.owner_host, self._addr.neighbour_host
        _, _, co, cn = self.level_maps(level - 1)
        return co, cn

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().pfm_gamg_destroy(h)
            except Exception:
                pass
            self._h = None


def hierarchy_for(matrix: LduMatrix, params: _lib.GamgParams) -> GamgHierarchy:
    addr = matrix.addressing
    cache = addr.__dict__.setdefault("_gamg", {})
    key = tuple(getattr(params, f) for f, _ in params._fields_)
    if key not in cache:
        cache[key] = GamgHierarchy(addr, params)
    return cache[key]


class GAMGSolver(LinearSolverBase):
    def __init__(self, tolerance: float = 1e-6, rel_tol: float = 0.01, max_iter: int = 100,
                 min_iter: int = 0, verbose: bool = False, n_pre_smooth: int = 2,
                 n_post_smooth: int = 2, max_levels: int = 10, min_cells_coarse: int = 10,
                 smoother: str = "PCG", agglomeration: str = "greedy",
                 coarse: str = "reference", n_coarsest: int = 50,
                 level_smoother: str = "Jacobi", precision: str = "double",
                 inner_rel_tol: float = 0.1, max_outer: int = 50, fp32_levels: int = 64) -> None:
        super().__init__(tolerance=tolerance, rel_tol=rel_tol, max_iter=max_iter,
                         min_iter=min_iter, verbose=verbose)
        self._n_pre_smooth = n_pre_smooth
        self._n_post_smooth = n_post_smooth
        self._max_levels = max_levels
        self._min_cells_coarse = min_cells_coarse
        self._smoother_type = smoother.upper()
        key = agglomeration.lower().replace("_", "").replace("-", "")
        if key not in _AGGLOM:
            raise ValueError(f"Unknown agglomeration '{agglomeration}'. Use 'greedy' or 'pair'.")
        if coarse.lower() not in _COARSE:
            raise ValueError(f"Unknown coarse operator '{coarse}'. Use 'reference' or 'galerkin'.")
        sk = level_smoother.lower().replace("_", "").replace("-", "")
        if sk not in _LEVEL_SMOOTHER:
            raise ValueError(f"Unknown level smoother '{level_smoother}'. "
                             "Use 'Jacobi' or 'GaussSeidel'.")
        self._params = _lib.GamgParams(max_levels, min_cells_coarse, n_pre_smooth, n_post_smooth,
                                       n_coarsest, _AGGLOM[key], _COARSE[coarse.lower()],
                                       _LEVEL_SMOOTHER[sk])
        if precision not in ("double", "mixed"):
            raise ValueError(f"Unknown precision '{precision}'. Use 'double' or 'mixed'.")
        self._precision = precision
        self._inner_rel_tol = float(inner_rel_tol)
        self._max_outer = int(max_outer)
        self._fp32_levels = int(fp32_levels)
        self._ws = KrylovWorkspace()

    def hierarchy(self, matrix: LduMatrix) -> GamgHierarchy:
        return hierarchy_for(matrix, self._params)

    def _solve(self, matrix: LduMatrix, source: torch.Tensor, x0: torch.Tensor,
               monitor: ResidualMonitor, max_iter: int) -> tuple[torch.Tensor, ConvergenceInfo]:
        dev, dt = matrix.device, matrix.dtype
        require_gpu(matrix.diag)
        b = source.to(device=dev, dtype=dt).contiguous()
        if b.shape[0] != matrix.n_cells:
            raise ValueError(f"source has {b.shape[0]} elements, expected {matrix.n_cells}")
        x = x0.to(device=dev, dtype=dt).clone().contiguous()
        max_iter = int(max_iter)
        if max_iter < 0:
            raise ValueError("max_iter must be >= 0")
        if matrix.n_cells == 0:
            monitor.load_history([])
            return x, monitor.build_info(False)
        hier = self.hierarchy(matrix)
        L = _lib.lib()
        prm = _lib.SolveParams(float(monitor.tolerance), float(monitor.rel_tol), max_iter,
                               int(monitor.min_iter), 0, 0)
        res = _lib.SolveResult()
        if self._precision == "mixed":
            return self._solve_mixed(matrix, hier, prm, b, x, monitor)
        work = self._ws.get(int(L.pfm_gamg_workspace_bytes(hier.handle, max_iter)), dev)
        hist = np.zeros(max_iter + 1)
        d, lo, up = matrix.coefficients()
        _lib.check(L.pfm_gamg_solve(hier.handle, C.byref(prm),
                                    _lib.ptr(d), _lib.ptr(lo), _lib.ptr(up), _lib.ptr(b),
                                    _lib.ptr(x), _lib.ptr(work), hist.ctypes.data, C.byref(res),
                                    current_stream(dev)), "pfm_gamg_solve")
        monitor.load_history(hist[: min(res.iterations, max_iter + 1)].tolist())
        info = monitor.build_info(converged=bool(res.converged))
        info.status = int(res.status)
        return x, info

    def _solve_mixed(self, matrix, hier, prm, b, x, monitor):
        L = _lib.lib()
        dev = matrix.device
        mo = self._max_outer
        work = self._ws.get(int(L.pfm_gamg_mixed_workspace_bytes(hier.handle, mo,
                                                                 self._fp32_levels)), dev)
        res = _lib.SolveResult()
        hist = np.zeros(mo + 1)
        inner = C.c_int32()
        d, lo, up = matrix.coefficients()
        _lib.check(L.pfm_gamg_mixed_solve(hier.handle, C.byref(prm), self._inner_rel_tol, mo,
                                          self._fp32_levels,
                                          _lib.ptr(d), _lib.ptr(lo), _lib.ptr(up),
                                          _lib.ptr(b), _lib.ptr(x), _lib.ptr(work),
                                          hist.ctypes.data, C.byref(res), C.byref(inner),
                                          current_stream(dev)), "pfm_gamg_mixed_solve")
        monitor.load_history(hist[: min(res.iterations, mo + 1)].tolist())
        info = monitor.build_info(converged=bool(res.converged))
        info.status = int(res.status)
        info.inner_iterations = int(inner.value)
        return x, info

    def pcg_solve(self, matrix: LduMatrix, b: torch.Tensor, x: torch.Tensor,
                  monitor: ResidualMonitor, max_iter: int, fp32_levels: int):
        """PCG preconditioned by one V-cycle of this solver's hierarchy
        (``pfm_gamg_pcg_solve``); x is updated in place."""
        hier = self.hierarchy(matrix)
        L = _lib.lib()
        dev = matrix.device
        work = self._ws.get(int(L.pfm_gamg_pcg_workspace_bytes(hier.handle, max_iter,
                                                               fp32_levels)), dev)
        prm = _lib.SolveParams(float(monitor.tolerance), float(monitor.rel_tol), max_iter,
                               int(monitor.min_iter), 0, 0)
        res = _lib.SolveResult()
        hist = np.zeros(max_iter + 1)
        d, lo, up = matrix.coefficients()
        _lib.check(L.pfm_gamg_pcg_solve(hier.handle, C.byref(prm), int(fp32_levels),
                                        _lib.ptr(d), _lib.ptr(lo), _lib.ptr(up), _lib.ptr(b),
                                        _lib.ptr(x), _lib.ptr(work), hist.ctypes.data,
                                        C.byref(res), current_stream(dev)), "pfm_gamg_pcg_solve")
        monitor.load_history(hist[: min(res.iterations, max_iter + 1)].tolist())
        info = monitor.build_info(converged=bool(res.converged))
        info.status = int(res.status)
        return x, info

    def coarse_values(self, matrix: LduMatrix, level: int):
        """(diag, lower, upper) of a coarse level as the solver builds them."""
        hier = self.hierarchy(matrix)
        dev = matrix.device
        L = _lib.lib()
        work = self._ws.get(int(L.pfm_gamg_workspace_bytes(hier.handle, 0)), dev)
        s = current_stream(dev)
        dd, ll, uu = matrix.coefficients()
        _lib.check(L.pfm_gamg_setup_values(hier.handle, _lib.ptr(dd), _lib.ptr(ll), _lib.ptr(uu),
                                           _lib.ptr(work), s), "pfm_gamg_setup_values")
        n, nf = hier.level_sizes()[level]
        d = torch.empty(n, dtype=torch.float64, device=dev)
        lo = torch.empty(nf, dtype=torch.float64, device=dev)
        up = torch.empty(nf, dtype=torch.float64, device=dev)
        _lib.check(L.pfm_gamg_level_values(hier.handle, _lib.ptr(work), level, _lib.ptr(d),
                                           _lib.ptr(lo), _lib.ptr(up), s), "pfm_gamg_level_values")
        return d, lo, up
