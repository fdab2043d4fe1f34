"""Host-side setup of libpfm (no GPU): GAMG aggregation and pair
agglomeration against the oracle and the reference's golden maps."""

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

from oracle import gamg as ogamg
from pyopenfoam_amd import _lib
from tests.helpers import case

G = Path(__file__).resolve().parent / "golden"


def native_aggregate(n, own, nei, target):
    o, q = _lib.host_i32(own), _lib.host_i32(nei)
    f2c = np.empty(n, np.int32)
    nc = C.c_int32()
    _lib.check(_lib.lib().pfm_gamg_aggregate(n, o.shape[0], o.ctypes.data, q.ctypes.data, target,
                                             f2c.ctypes.data, C.byref(nc)))
    return nc.value, f2c


def native_pairs(n, own, nei, n_levels):
    o, q = _lib.host_i32(own), _lib.host_i32(nei)
    maps = np.empty((n_levels, n), np.int32)
    counts = np.empty(n_levels, np.int32)
    made = C.c_int32()
    _lib.check(_lib.lib().pfm_pair_agglomerate(n, o.shape[0], o.ctypes.data, q.ctypes.data,
                                               n_levels, maps.ctypes.data, counts.ctypes.data,
                                               C.byref(made)))
    return [(int(counts[k]), maps[k]) for k in range(made.value)]


@pytest.mark.parametrize("name", ["sys_4", "sys_6"])
def test_aggregation_matches_golden(name):
    g = dict(np.load(G / f"{name}.npz"))
    nfi = g["neighbour"].shape[0]
    n = int(g["n_cells"])
    nc, f2c = native_aggregate(n, g["owner"][:nfi], g["neighbour"], max(n // 2, 4))
    assert nc == int(g["gamg_ncoarse"])
    assert np.array_equal(f2c, g["gamg_f2c"])
    lv = native_pairs(n, g["owner"][:nfi], g["neighbour"], 5)
    assert len(lv) == int(g["pair_nlevels"])
    for k, (cnt, m) in enumerate(lv):
        assert np.array_equal(m, g[f"pair_map{k}"])
        assert cnt == int(g[f"pair_n{k}"][1])


@pytest.mark.parametrize("n,jit", [(8, 0.0), (12, 0.2), (20, 0.2)])
def test_aggregation_matches_oracle(n, jit):
    c = case(n, jit, 1.0)
    for target in (c.n_cells // 2, c.n_cells // 9, 10):
        nc, f2c = native_aggregate(c.n_cells, c.owner, c.neighbour, target)
        onc, of2c = ogamg.aggregate(c.n_cells, c.owner, c.neighbour, target)
        assert nc == onc and np.array_equal(f2c, of2c)
    lv = native_pairs(c.n_cells, c.owner, c.neighbour, 8)
    ol = ogamg.pair_levels(c.n_cells, c.owner, c.neighbour, n_levels=8)
    assert len(lv) == len(ol)
    for (cnt, m), o in zip(lv, ol):
        assert cnt == o["n_cells_coarse"] and np.array_equal(m, o["cell_map"])


def test_pair_agglomeration_chain_known_answers():
    # reference tests/unit/fv_agglomeration/test_agglomeration.py:14-47
    lv = native_pairs(16, np.arange(15), np.arange(1, 16), 3)
    assert len(lv) > 0 and lv[0][0] < 16
    assert native_pairs(4, np.array([0, 1, 2]), np.array([1, 2, 3]), 5) == []
    # chain of 8: faces pair (0,1), (2,3), (4,5), (6,7) -> 4 cells, then 2
    lv = native_pairs(8, np.arange(7), np.arange(1, 8), 5)
    assert lv[0][0] == 4 and lv[0][1].tolist() == [0, 0, 1, 1, 2, 2, 3, 3]
