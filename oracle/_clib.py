"""ctypes loader of the oracle's C restatement (oracle/_build/liboracle.so)."""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_SO = _HERE / "_build" / "liboracle.so"
_lib = None

P = C.c_void_p
I64 = C.c_int64
D = C.c_double

_SIGS = {
    "orc_dic_build": (C.c_int, [I64, I64, P, P, P, P, P]),
    "orc_dilu_build": (C.c_int, [I64, I64, P, P, P, P, P, P]),
    "orc_dic_apply_full": (C.c_int, [I64, I64, P, P, P, P, P, P]),
    "orc_dilu_apply_full": (C.c_int, [I64, I64, P, P, P, P, P, P, P]),
    "orc_gauss_seidel": (C.c_int, [I64, I64, P, P, P, P, P, P, D, I64, P]),
    "orc_smooth_gs": (C.c_int, [I64, I64, P, P, P, P, P, P, I64, P]),
    "orc_gamg_aggregate": (I64, [I64, I64, P, P, I64, P]),
    "orc_gamg_coarse_faces": (I64, [I64, P, P, P, P, P, P, P, P, P, P]),
    "orc_pair_level": (I64, [I64, I64, I64, P, P, P, P]),
    "orc_norm3": (C.c_int, [I64, P, P]),
}


def build() -> Path:
    r = subprocess.run(["make", "-s", "-C", str(_HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return _SO


def clib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            build()
        h = C.CDLL(str(_SO))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.int64)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.float64)


def p(a: np.ndarray) -> int:
    return a.ctypes.data
