"""CPU-side checks of the gfx950 library: the C ABI loads and exports every
symbol include/loadsg_amd.h declares, and the host-side grid flattening (no GPU
calls) reproduces the reference's index sets and the closed-form sizes."""
import math
import os
import re

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import Oracle, regular_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "loadsg_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\*?(lsg_[a-z_0-9]+)\s*\(", src, re.M)))


def test_library_exports_declared_symbols():
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


@pytest.mark.parametrize("d,level", [(1, 6), (2, 6), (3, 4), (4, 3), (5, 3), (6, 2)])
def test_regular_grid_slot_order_matches_reference_order(d, level):
    g = L.AdaptiveSparseGrid.regular(d, level)
    idx, act = g.slot_order()
    ridx, ract = regular_grid(d, level)
    assert np.array_equal(idx, ridx)
    assert np.array_equal(act, ract)
    n = sum(2 ** k * math.comb(k + d - 1, d - 1) for k in range(level + 1))
    assert g.size == n


def test_device_order_is_depth_blocked_permutation():
    g = L.AdaptiveSparseGrid.regular(3, 4)
    idx, _ = g.slot_order()
    dev = g.device_order()
    assert np.array_equal(np.sort(dev), np.arange(g.size))
    depth = np.floor(np.log2(idx)).astype(int)
    # positions of one depth block are contiguous
    for T in np.unique(depth, axis=0):
        pos = np.sort(dev[np.all(depth == T, axis=1)])
        assert pos[-1] - pos[0] + 1 == pos.size


def test_grid_from_golden_index_table():
    gold = os.path.join(ROOT, "tests", "golden", "d4_l3_cfg2")
    idx = np.load(os.path.join(gold, "idx.npy"))
    act = np.load(os.path.join(gold, "active.npy"))
    g = L.AdaptiveSparseGrid(idx, act)
    i2, a2 = g.slot_order()
    assert np.array_equal(i2, idx) and np.array_equal(a2, act)
    o = Oracle(idx, act)
    assert g.stats()["multilevel_size"] == o.multilevel_size()


@pytest.mark.parametrize("d,level,stream", [(2, 6, 64302), (5, 7, 41792143)])
def test_streaming_model_value_count(d, level, stream):
    """Streaming-model values per matvec (every stage reads/writes its values once)."""
    g = L.AdaptiveSparseGrid.regular(d, level)
    s = g.stats()
    assert s["stream_values"] == stream
    assert s["max_box"] <= 16384


def test_adaptive_grid_takes_the_sparse_path():
    """A locally adaptive grid with hanging nodes (the reference's completeDirichletGrid
    output) is accepted: depth blocks are point lists, G_T the stored points of depth <= T
    (multilevel size equal to the oracle's), the device order a permutation."""
    from oracle_ctypes import Oracle
    gold = os.path.join(ROOT, "tests", "golden", "d3_adaptive_cfg3")
    idx = np.load(os.path.join(gold, "idx.npy"))
    act = np.load(os.path.join(gold, "active.npy"))
    g = L.AdaptiveSparseGrid(idx, act)
    assert g.size == idx.shape[0]
    st = g.stats()
    assert st["multilevel_size"] == Oracle(idx, act).multilevel_size()
    dev = g.device_order()
    assert sorted(dev.tolist()) == list(range(g.size))
    i2, a2 = g.slot_order()
    assert np.array_equal(i2, idx) and np.array_equal(a2, act)


@pytest.mark.parametrize("level", [7, 8, 9, 10])
def test_refined_cfg3_grids_build(level):
    """The config-3 style grids refined towards the origin (the reference's own refinement
    and completeDirichletGrid output, start level 4, sum level 7..10) build on the sparse
    path with the oracle's multilevel size."""
    from oracle_ctypes import Oracle
    z = np.load(os.path.join(ROOT, "tests", "golden", "grids", "cfg3_start4_level%d.npz" % level))
    g = L.AdaptiveSparseGrid(z["idx"], z["active"])
    assert g.size == z["idx"].shape[0] and (z["active"] == 0).any()
    assert g.stats()["multilevel_size"] == Oracle(z["idx"], z["active"]).multilevel_size()


def test_invalid_input_errors():
    with pytest.raises(L.LoadsgError):
        L.AdaptiveSparseGrid(np.array([[0, 2]], dtype=np.int32))        # boundary point
    with pytest.raises(L.LoadsgError):
        L.AdaptiveSparseGrid(np.array([[2, 3]], dtype=np.int32))        # index 3 is outside (0,1)
    with pytest.raises(L.LoadsgError):
        L.AdaptiveSparseGrid(np.array([[2, 2], [2, 2]], dtype=np.int32))  # duplicate


def test_capacity_and_structure_errors():
    """Grids the path cannot take return a status instead of failing later: a 1D depth
    above 15 (MaximumDepth, abbrevi.h:12), a depth list that is not downward closed."""
    with pytest.raises(L.LoadsgError, match="error -2"):
        L.AdaptiveSparseGrid(np.array([[2 ** 16, 2]], dtype=np.int32))      # 1D depth 16
    # depth (2,1) without (1,1): point (4,2) alone
    with pytest.raises(L.LoadsgError, match="error -2"):
        L.AdaptiveSparseGrid(np.array([[4, 2]], dtype=np.int32))


def test_cells_of_regular_grid():
    """lsg_grid_cells (host only): sum over depths T of prod 2^T_i cells of depth T+1 on a
    regular grid, centres inside (0,1), grouped by depth in depth-list order."""
    for d, level in ((2, 4), (3, 3)):
        g = L.AdaptiveSparseGrid.regular(d, level)
        x, t = g.cells()
        n = sum(2 ** k * math.comb(k + d - 1, d - 1) for k in range(level + 1))
        assert x.shape == (2 ** d * n, d)
        assert np.all((x > 0) & (x < 1))
        assert np.all(np.diff(t) >= 0)
