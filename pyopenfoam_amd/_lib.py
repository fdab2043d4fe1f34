"""ctypes binding of libpfm.so (the C ABI declared in include/pfm.h).

The library is loaded from the package directory.  There is no fallback: if
it is missing, every compute entry point raises ``RuntimeError`` telling the
caller to build it (``python -c 'import __graft_entry__ as g; g.build()'``).
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import numpy as np

__all__ = ["lib", "check", "PfmError", "SolveParams", "SolveResult", "LduInfo", "GamgParams",
           "library_path",
           "PFM_OK", "PFM_ERR_ARG", "PFM_ERR_HIP", "PFM_SOLVE_CONVERGED", "PFM_SOLVE_MAXITER",
           "PFM_SOLVE_BREAKDOWN", "PFM_PRECOND_NONE", "PFM_PRECOND_DIAG", "EXPORTED_SYMBOLS"]

PFM_OK, PFM_ERR_ARG, PFM_ERR_HIP, PFM_ERR_STATE, PFM_ERR_NOMEM = 0, 1, 2, 3, 4
PFM_SOLVE_RUNNING, PFM_SOLVE_CONVERGED, PFM_SOLVE_MAXITER, PFM_SOLVE_BREAKDOWN = 0, 1, 2, 3
PFM_PRECOND_NONE, PFM_PRECOND_DIAG = 0, 1

library_path = Path(__file__).resolve().parent / "libpfm.so"


class PfmError(RuntimeError):
    pass


class SolveParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("rel_tol", C.c_double), ("max_iter", C.c_int32),
                ("min_iter", C.c_int32), ("precond", C.c_int32), ("reserved", C.c_int32)]


class SolveResult(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("status", C.c_int32),
                ("reserved", C.c_int32), ("initial_residual", C.c_double),
                ("final_residual", C.c_double)]


class GamgParams(C.Structure):
    _fields_ = [("max_levels", C.c_int32), ("min_cells_coarse", C.c_int32), ("n_pre", C.c_int32),
                ("n_post", C.c_int32), ("n_coarsest", C.c_int32), ("agglomeration", C.c_int32),
                ("coarse", C.c_int32), ("smoother", C.c_int32)]


PFM_AGGLOM_GREEDY, PFM_AGGLOM_PAIR = 0, 1
PFM_SMOOTHER_JACOBI, PFM_SMOOTHER_GAUSS_SEIDEL = 0, 1
PFM_SMOOTH_GAUSS_SEIDEL, PFM_SMOOTH_DIC, PFM_SMOOTH_DILU = 0, 1, 2
PFM_COARSE_REFERENCE, PFM_COARSE_GALERKIN = 0, 1


class LduInfo(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("n_faces", C.c_int32), ("n_slices", C.c_int32),
                ("n_dic_levels", C.c_int32), ("n_dilu_levels", C.c_int32),
                ("max_lower", C.c_int32), ("max_upper", C.c_int32), ("device", C.c_int32),
                ("n_slots", C.c_int64)]


P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
INT = C.c_int

# name -> (restype, argtypes)
_SIGS = {
    "pfm_last_error": (C.c_char_p, []),
    "pfm_abi_version": (INT, []),
    "pfm_rcm": (INT, [I32, I32, P, P, P]),
    "pfm_parse_ascii_points": (INT, [C.c_char_p, I64, I64, P, P]),
    "pfm_parse_ascii_faces": (INT, [C.c_char_p, I64, I64, I64, P, P, P, P]),
    "pfm_parse_ascii_labels": (INT, [C.c_char_p, I64, I64, P, P]),
    "pfm_cell_faces": (INT, [I32, I32, I32, P, P, P, P, P, P]),
    "pfm_relabel_faces": (INT, [I32, I32, P, P, P, P, P, P, P]),
    "pfm_face_weights": (INT, [I32, I32, P, P, P, P, P, P]),
    "pfm_pressure_assemble": (INT, [P, I32, P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "pfm_correct_face_flux": (INT, [P, I32, P, P, P, P, P, P, P, P]),
    "pfm_correct_velocity": (INT, [P, I32, P, P, P, P, P, P, P, P, C.c_double, P, P]),
    "pfm_adjust_phi": (INT, [I32, I32, P, INT, P, P, P, P]),
    "pfm_mesh_geometry": (INT, [I32, P, I32, P, P, I32, P, P, I32, P, P, I32, P, P, P, P, P, P, P]),
    "pfm_dic_levels": (INT, [I32, I32, P, P, P, P]),
    "pfm_dilu_levels": (INT, [I32, I32, P, P, P, P]),
    "pfm_gs_levels": (INT, [I32, I32, P, P, P, P, P, P]),
    "pfm_gamg_aggregate": (INT, [I32, I32, P, P, I32, P, P]),
    "pfm_pair_agglomerate": (INT, [I32, I32, P, P, I32, P, P, P]),
    "pfm_ldu_create": (INT, [I32, I32, P, P, C.POINTER(P)]),
    "pfm_ldu_destroy": (INT, [P]),
    "pfm_ldu_get_info": (INT, [P, C.POINTER(LduInfo)]),
    "pfm_ldu_pack": (INT, [P, INT, P, P, P, P]),
    "pfm_ldu_pack_f32": (INT, [P, INT, P, P, P, P]),
    "pfm_spmv": (INT, [P, P, P, P, P, P]),
    "pfm_spmv_f32": (INT, [P, P, P, P, P, P]),
    "pfm_bandwidth_probe": (INT, [INT, P, P, I64, P]),
    "pfm_spmv_multi": (INT, [P, P, P, I32, P, P, P]),
    "pfm_relax": (INT, [I32, C.c_double, P, P, P, P, P, P]),
    "pfm_bandwidth_probe_out_bytes": (I64, []),
    "pfm_spmv_ldu": (INT, [P, INT, P, P, P, P, P, P]),
    "pfm_residual": (INT, [P, P, P, P, P, P, P]),
    "pfm_ldu_set_progress": (INT, [P, P, P, C.c_double]),
    "pfm_gs_smooth": (INT, [P, P, P, P, P, I32, C.c_double, P]),
    "pfm_precond_apply_full": (INT, [P, INT, P, P, P, P, P, P]),
    "pfm_diagonal_solve": (INT, [P, P, P, P, P, P, P]),
    "pfm_smooth_workspace_bytes": (I64, [P, I32]),
    "pfm_smooth_solve": (INT, [P, C.POINTER(SolveParams), INT, I32, C.c_double, P, P, P, P, P, P,
                               P, P, P, C.POINTER(SolveResult), P]),
    "pfm_jacobi_smooth": (INT, [P, P, P, P, P, I32, C.c_double, P, P]),
    "pfm_dic_factor": (INT, [P, P, P, P, P]),
    "pfm_dilu_factor": (INT, [P, INT, P, P, P, P, P]),
    "pfm_pcg_workspace_bytes": (I64, [P, I32]),
    "pfm_pbicgstab_workspace_bytes": (I64, [P, I32]),
    "pfm_pcg_solve": (INT, [P, C.POINTER(SolveParams), P, P, P, P, P, P, P, C.POINTER(SolveResult), P]),
    "pfm_pbicgstab_solve": (INT, [P, C.POINTER(SolveParams), P, P, P, P, P, P, P,
                                  C.POINTER(SolveResult), P]),
    "pfm_pcg_noalloc_max_bytes": (I64, []),
    "pfm_pcg_solve_noalloc": (INT, [P, C.POINTER(SolveParams), P, P, P, P, P, P,
                                    C.POINTER(SolveResult), P]),
    "pfm_pbicg_workspace_bytes": (I64, [P, I32]),
    "pfm_pbicg_solve": (INT, [P, C.POINTER(SolveParams), P, P, P, P, P, P, P, P,
                              C.POINTER(SolveResult), P]),
    "pfm_pcg_mixed_workspace_bytes": (I64, [P, I32]),
    "pfm_pcg_mixed_solve": (INT, [P, C.POINTER(SolveParams), C.c_double, I32, P, P, P, P, P, P, P,
                                  P, C.POINTER(SolveResult), P, P]),
    "pfm_batch_create": (INT, [I32, P, P, P, P, I32, C.POINTER(P)]),
    "pfm_batch_destroy": (INT, [P]),
    "pfm_batch_info": (INT, [P, P, P, P, P]),
    "pfm_batch_workspace_bytes": (I64, [P]),
    "pfm_batch_pcg_solve": (INT, [P, C.POINTER(SolveParams), INT, P, P, P, P, P, P, I32, P, P, P,
                                  P, P]),
    "pfm_gamg_create": (INT, [P, C.POINTER(GamgParams), C.POINTER(P)]),
    "pfm_gamg_destroy": (INT, [P]),
    "pfm_gamg_num_levels": (INT, [P]),
    "pfm_gamg_level_info": (INT, [P, I32, P, P]),
    "pfm_gamg_level_maps": (INT, [P, I32, P, P, P, P]),
    "pfm_gamg_workspace_bytes": (I64, [P, I32]),
    "pfm_gamg_setup_values": (INT, [P, P, P, P, P, P]),
    "pfm_gamg_level_values": (INT, [P, P, I32, P, P, P, P]),
    "pfm_gamg_solve": (INT, [P, C.POINTER(SolveParams), P, P, P, P, P, P, P, C.POINTER(SolveResult), P]),
    "pfm_gamg_mixed_workspace_bytes": (I64, [P, I32, I32]),
    "pfm_gamg_pcg_workspace_bytes": (I64, [P, I32, I32]),
    "pfm_gamg_pcg_solve": (INT, [P, C.POINTER(SolveParams), I32, P, P, P, P, P, P, P,
                                 C.POINTER(SolveResult), P]),
    "pfm_gamg_mixed_solve": (INT, [P, C.POINTER(SolveParams), C.c_double, I32, I32, P, P, P, P, P,
                                   P, P, C.POINTER(SolveResult), P, P]),
    "pfm_dirichlet_create": (INT, [P, I32, P, I32, P, C.POINTER(P)]),
    "pfm_dirichlet_destroy": (INT, [P]),
    "pfm_laplacian_assemble": (INT, [P, P, P, P, P, P, P, P, P, P, P, P, P]),
    "pfm_adjoint_ldu_grads": (INT, [P, P, P, P, P, P, P]),
    "pfm_laplacian_backward": (INT, [P, P, P, P, P, P, P, I32, P, P, P, P, P, P]),
    "pfm_adjoint_gamma_grad": (INT, [P, P, P, P, P, P, I32, P, P, P, P, P, P]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)

# pfm_progress_fn: void (*)(int32_t iterations, void *user)
PROGRESS_FN = C.CFUNCTYPE(None, C.c_int32, C.c_void_p)

_lock = threading.Lock()
_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not library_path.exists():
                raise PfmError(
                    f"{library_path} is missing: build the HIP extension first "
                    "(python -c 'import __graft_entry__ as g; g.build()')")
            handle = C.CDLL(str(library_path))
            for name, (res, args) in _SIGS.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
    return _lib


def check(status: int, what: str = "") -> None:
    """Map a libpfm status to the reference's exception types."""
    if status == PFM_OK:
        return
    msg = lib().pfm_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if status == PFM_ERR_ARG:
        raise ValueError(text)
    raise PfmError(text)


def host_i32(a) -> np.ndarray:
    """Contiguous int32 host copy of an index array (numpy or torch)."""
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    a = np.asarray(a)
    if a.size and (a.min() < 0 or a.max() > np.iinfo(np.int32).max):
        raise ValueError("index out of int32 range")
    return np.ascontiguousarray(a, dtype=np.int32)


def ptr(a) -> int | None:
    """Raw address of a numpy array or torch tensor (None for None)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
