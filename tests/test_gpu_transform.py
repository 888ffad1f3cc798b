"""GPU parity of the basis transformations and energy norms (through the C ABI):
lsg_hierarchize / lsg_dehierarchize / lsg_prew_by_nodal / lsg_nodal_by_prew /
lsg_energy_norm against the compiled reference's golden vectors
(CalcHierarchicalBasis, calcPrewByNodal, calcNodalByPrew, H1/L2::getValue) on regular
and adaptive grids (hanging nodes; one or two connected components per depth), against
the CPU oracle up to d=6 and on the config-3 refined grids up to level 10, and through
size-independent properties at d=5 level 7 and on the level-11 refined grid.

Tolerances: 1e-12 relative (max-norm) against the reference and the oracle (the
device sums the same terms in another order and solves the line systems by an
unpivoted LU); round trips within 1e-12; batched instances bit-identical to single
ones.
"""
import json
import math
import os

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import Oracle, regular_grid

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
REGULAR = sorted(c for c in CASES if CASES[c]["argv"][0] == "regular")
ADAPTIVE = sorted(c for c in CASES if CASES[c]["argv"][0].startswith("adaptive"))
GRIDS = os.path.join(GOLD, "grids")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def load(case):
    g = os.path.join(GOLD, case)
    return {f[:-4]: np.load(os.path.join(g, f)) for f in os.listdir(g) if f.endswith(".npy")}


def coords(idx):
    depth = np.floor(np.log2(idx)).astype(int)
    return (2 * (idx - 2 ** depth) + 1) / 2.0 ** depth


@pytest.mark.parametrize("case", REGULAR + ADAPTIVE)
def test_transforms_match_reference_golden(case):
    gv = load(case)
    print("transform golden", case, flush=True)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    u = gv["tr_u"]
    assert relmax(L.CalcHierarchicalBasis(u, op), gv["tr_hier"]) < 1e-12
    assert relmax(L.calcPrewByNodal(u, op), gv["tr_prew"]) < 1e-12
    assert relmax(L.calcNodalByPrew(gv["x"], op), gv["tr_nodal_of_x"]) < 1e-12
    d = CASES[case]["d"]
    f = d * math.pi ** 2 * np.prod(np.sin(math.pi * coords(gv["idx"])), axis=1) * gv["active"]
    assert relmax(L.calcPrewByNodal(f, op), gv["f_prew"]) < 1e-12
    h1 = L.H1(grid).getValue(u)
    l2 = L.L2(grid).getValue(u)
    assert abs(h1 - gv["tr_norms"][0]) < 1e-12 * gv["tr_norms"][0]
    assert abs(l2 - gv["tr_norms"][1]) < 1e-12 * gv["tr_norms"][1]


@pytest.mark.parametrize("d,level", [(1, 10), (2, 8), (3, 6), (4, 5), (5, 4), (6, 3)])
def test_transforms_match_oracle(d, level):
    print("transform oracle d=%d L=%d" % (d, level), flush=True)
    idx, act = regular_grid(d, level)
    o = Oracle(idx, act)
    grid = L.AdaptiveSparseGrid(idx, act)
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    rng = np.random.default_rng(100 * d + level)
    u = rng.standard_normal(idx.shape[0])
    p = rng.standard_normal(idx.shape[0])
    assert relmax(op.hierarchize(u), o.transform(o.HIER, u)) < 1e-12
    assert relmax(op.dehierarchize(u), o.transform(o.DEHIER, u)) < 1e-12
    assert relmax(op.prew_by_nodal(u), o.transform(o.PREW_BY_NODAL, u)) < 1e-12
    assert relmax(op.nodal_by_prew(p), o.transform(o.NODAL_BY_PREW, p)) < 1e-12


def refined_grid(level):
    z = np.load(os.path.join(GRIDS, "cfg3_start4_level%d.npz" % level))
    return z["idx"], z["active"]


@pytest.mark.parametrize("level", [8, 10])
def test_adaptive_transforms_match_oracle(level):
    """Config-3 refined grids (start level 4, refinement near the origin): every
    transform against the oracle for random values at every point, hanging ones
    included (they are read as poles and line values but never written)."""
    print("adaptive transform oracle level %d" % level, flush=True)
    idx, act = refined_grid(level)
    o = Oracle(idx, act)
    grid = L.AdaptiveSparseGrid(idx, act)
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    rng = np.random.default_rng(level)
    u = rng.standard_normal(idx.shape[0])
    p = rng.standard_normal(idx.shape[0])
    assert relmax(op.hierarchize(u), o.transform(o.HIER, u)) < 1e-12
    assert relmax(op.dehierarchize(u), o.transform(o.DEHIER, u)) < 1e-12
    assert relmax(op.prew_by_nodal(u), o.transform(o.PREW_BY_NODAL, u)) < 1e-12
    assert relmax(op.nodal_by_prew(p), o.transform(o.NODAL_BY_PREW, p)) < 1e-12


def test_adaptive_batch_and_global_boxes(monkeypatch):
    """Two-component grid: a batch of three instances (boxes of several instances share a
    workgroup) gives each single instance's result bit for bit, and so do boxes staged in
    global memory instead of LDS (LSG_TR_LDS_BYTES=0)."""
    import torch
    print("adaptive transform batch / global boxes", flush=True)
    gv = load("d3_adaptive2_cfg3")
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    n = grid.size
    rng = np.random.default_rng(7)
    U = rng.standard_normal((3, n))
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    singles = [(op.prew_by_nodal(U[b]), op.nodal_by_prew(U[b])) for b in range(3)]
    opb = L.MatrixVectorHomogen(grid, L.Poisson(), 64, batch=3)
    dev = grid.device_order()
    Ud = np.zeros_like(U)
    Ud[:, dev] = U
    tin = torch.tensor(Ud, dtype=torch.float64, device="cuda")
    for lds in (None, "0"):
        if lds is not None:
            monkeypatch.setenv("LSG_TR_LDS_BYTES", lds)
        tp = torch.zeros_like(tin)
        tn = torch.zeros_like(tin)
        opb.prew_by_nodal(tin, tp)
        opb.nodal_by_prew(tin, tn)
        P = tp.cpu().numpy()[:, dev]
        N = tn.cpu().numpy()[:, dev]
        for b in range(3):
            assert np.array_equal(P[b], singles[b][0]), (lds, b)
            assert np.array_equal(N[b], singles[b][1]), (lds, b)


def test_adaptive_level11_round_trip():
    """Refined grid of level 11 (N = 65,943; boxes up to 14,415 values, beyond 64 KB of
    LDS): nodal -> prewavelets -> nodal is the identity on the active points, hanging
    points are never written, and H1 = sqrt(p^T A p)."""
    print("adaptive transform level 11 round trip", flush=True)
    idx, act = refined_grid(11)
    grid = L.AdaptiveSparseGrid(idx, act)
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    a = act.astype(bool)
    u = np.random.default_rng(11).standard_normal(grid.size) * a
    p = op.prew_by_nodal(u)
    assert not p[~a].any()
    back = op.nodal_by_prew(p)
    assert np.abs(back - u)[a].max() < 1e-12 * np.abs(u).max()
    assert not back[~a].any()
    Ap = op.multiplication(p)
    assert abs(op.energy_norm(u) - math.sqrt(p @ Ap)) < 1e-12 * math.sqrt(p @ Ap)


def test_full_size_round_trips_and_batch():
    """d=5 level 7 (N = 61,183): nodal -> prewavelets -> nodal and nodal -> hierarchical
    -> nodal are identities; transforms are linear; a batch of three instances gives
    each single instance's result bit for bit; the H1 norm equals sqrt(p^T A p)."""
    import torch
    print("transform d=5 L=7 round trips", flush=True)
    grid = L.AdaptiveSparseGrid.regular(5, 7)
    n = grid.size
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    rng = np.random.default_rng(5)
    u = rng.standard_normal(n)
    v = rng.standard_normal(n)
    p = op.prew_by_nodal(u)
    assert relmax(op.nodal_by_prew(p), u) < 1e-12
    assert relmax(op.dehierarchize(op.hierarchize(u)), u) < 1e-13
    pv = op.prew_by_nodal(v)
    assert relmax(op.prew_by_nodal(2.0 * u - 3.0 * v), 2.0 * p - 3.0 * pv) < 1e-12
    # energy norm = sqrt(p . A p)
    Ap = op.multiplication(p)
    assert abs(op.energy_norm(u) - math.sqrt(p @ Ap)) < 1e-12 * math.sqrt(p @ Ap)
    # batch of 3 on device vectors (device order) == single instances
    opb = L.MatrixVectorHomogen(grid, L.Poisson(), 64, batch=3)
    dev = grid.device_order()
    U = np.stack([u, v, u - v])
    Ud = np.zeros_like(U)
    Ud[:, dev] = U
    tin = torch.tensor(Ud, dtype=torch.float64, device="cuda")
    tout = torch.zeros_like(tin)
    opb.prew_by_nodal(tin, tout)
    P = tout.cpu().numpy()[:, dev]
    for b in range(3):
        single = op.prew_by_nodal(U[b])
        assert np.array_equal(P[b], single)
    norms = opb.energy_norm(tin)
    for b in range(3):
        assert norms[b] == op.energy_norm(U[b])


def test_d6_l8_round_trip():
    """Largest config-4 shape (N = 471,041; depth boxes up to 3,969 values)."""
    print("transform d=6 L=8 round trip", flush=True)
    grid = L.AdaptiveSparseGrid.regular(6, 8)
    op = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    u = np.random.default_rng(68).standard_normal(grid.size)
    assert relmax(op.nodal_by_prew(op.prew_by_nodal(u)), u) < 1e-12
    assert relmax(op.dehierarchize(op.hierarchize(u)), u) < 1e-13


def test_poisson_driver_known_answers():
    """The reference driver examples_doxygen/poisson_regular.cpp (d=3, level 7) entirely on
    the device: f_h = M prew(f_I), CG to eps 1e-10 with the reference stop rule,
    u_h = calcNodalByPrew(x), and the errors of u - u_h in the reference's norms, against
    the values that driver prints (six significant digits)."""
    from test_oracle_transform import POISSON_D3_L7 as K
    print("poisson driver d=3 L=7", flush=True)
    d = 3
    grid = L.AdaptiveSparseGrid.regular(d, 7)
    idx, _ = grid.slot_order()
    u = np.prod(np.sin(math.pi * coords(idx)), axis=1)
    poisson = L.MatrixVectorHomogen(grid, L.Poisson(), 64)
    mass = L.MatrixVectorHomogen(grid, L.HelmHoltz(), 64)
    f = mass.multiplication(L.calcPrewByNodal(d * math.pi ** 2 * u, poisson))
    x = np.zeros(grid.size)
    info = L.CG.solveHomogen(1e-10, x, f, poisson)
    e = u - L.calcNodalByPrew(x, poisson)
    l2 = L.L2(grid).getValue(e)
    h1 = L.H1(grid).getValue(e)
    assert info["iterations"] == K["iterations"]
    assert abs(np.abs(e).max() - K["linf"]) < 1e-5 * K["linf"]
    assert abs(l2 - K["l2"]) < 1e-5 * K["l2"]
    assert abs(l2 + h1 - K["h1_plus_l2"]) < 1e-5 * K["h1_plus_l2"]


@pytest.mark.parametrize("where", ["regular", "adaptive"])
@pytest.mark.parametrize("prec", [64, 32])
def test_vector_norms(where, prec):
    """L_infty over all stored points, L_2 over the active ones (norms.h:12-100)."""
    print("vector norms", where, prec, flush=True)
    if where == "regular":
        grid = L.AdaptiveSparseGrid.regular(4, 5)
        act = np.ones(grid.size, dtype=bool)
    else:
        gv = load("d3_adaptive_cfg3")
        grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
        act = gv["active"].astype(bool)
        assert not act.all()
    op = L.MatrixVectorHomogen(grid, L.Poisson(), prec)
    x = np.random.default_rng(4).standard_normal(grid.size)
    x[np.argmax(~act) if not act.all() else 0] = 7.5     # largest entry (a hanging node when there is one)
    if prec == 32:
        x = x.astype(np.float32).astype(np.float64)
    linf, l2 = op.vector_norms(x)
    assert linf == np.max(np.abs(x))
    assert abs(l2 - np.sqrt(np.sum(x[act] ** 2) / act.sum())) < 1e-13 * l2
    assert L.L_infty(x, op) == linf and L.L_2(x, op) == l2


def test_unsupported_inputs_raise():
    print("transform errors", flush=True)
    reg = L.AdaptiveSparseGrid.regular(3, 4)
    op32 = L.MatrixVectorHomogen(reg, L.Poisson(), 32)
    with pytest.raises(L.LoadsgError):
        op32.hierarchize(np.ones(reg.size))
