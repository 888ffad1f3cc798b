"""General coefficients across the C ABI (LSG_COEF_CELL / LSG_COEF_FUNC): a(x) and c(x)
sampled on the host at the centre of every cell of depth T+1, the value a
StencilTemplate::integration plug-in (stencils/Stencil.h:740) evaluates per cell.

Checked against the compiled reference's golden vectors for the non-separable pair
a = 1 + x1 x2 + sin(pi (x1 + x2)), c = 1 + sin(pi x1 x2) / 2 (the harness's "gen"
plug-in, oracle/ref_plugin.hpp) on regular and adaptive grids, and against the CPU
oracle.  Tolerances: FP64 1e-12 relative (max-norm), FP32 2e-5, iteration counts exact.
"""
import json
import os

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import COEF_EXPSIN, COEF_GEN, Oracle, gen_coef, regular_grid
from test_gpu_parity import load, relmax

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
GEN = sorted(c for c in CASES if CASES[c]["argv"][-1] == "gen")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


def cell_values(grid):
    x, _ = grid.cells()
    return gen_coef(x)


@pytest.mark.parametrize("case", GEN)
def test_cell_coefficient_golden(case):
    """Matvec (FP64 and FP32), the diagonal and CG (reference stop rule) against the
    reference's own results with the same plug-in; hanging nodes stay zero."""
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    a, c = cell_values(grid)
    st = L.CellCoefficient(a, c)
    op = L.MatrixVectorHomogen(grid, st)
    y = op.multiplication(gv["x"])
    assert relmax(y, gv["Ax_var"]) < 1e-12
    assert np.all(y[gv["active"] == 0] == 0.0)
    assert relmax(L.MatrixVectorHomogen(grid, st, precision=32).multiplication(gv["x"]), gv["Ax_var"]) < 2e-5
    act = gv["active"] == 1
    o = Oracle(gv["idx"], gv["active"]).set_coef(kind=COEF_GEN)
    z = op.diagonal()
    assert relmax(z[act], o.precond()[act]) < 1e-12
    if act.all():
        assert relmax(z, gv["diag_var"]) < 1e-11
    x = np.zeros(grid.size)
    info = L.CG.solveHomogen(1e-10, x, gv["f_rhs"], op, stop_rule=L.STOP_REF_ABS, maxit=3000)
    xo, ito, _ = o.cg(gv["f_rhs"], eps=1e-10)
    assert info["iterations"] == ito
    assert relmax(x, xo) < 1e-10
    if act.all():       # on adaptive grids the reference's diagonal carries stale hanging-node values
        assert info["iterations"] == int(gv["cg_var_iters"][0])
        assert relmax(x, gv["cg_var_x"]) < 1e-10


def test_function_route_equals_cell_route():
    """Host callbacks (LSG_COEF_FUNC) and per-cell arrays (LSG_COEF_CELL) give the same
    operator bit for bit; a constant cell coefficient reproduces the Poisson / HelmHoltz
    operators of the dense path within rounding."""
    gv = load("d3_adaptive_gen")
    for idx, act in ((gv["idx"], gv["active"]), regular_grid(3, 4)):
        grid = L.AdaptiveSparseGrid(idx, act)
        a, c = cell_values(grid)
        x = np.random.default_rng(4).standard_normal(grid.size) * act
        y_cell = L.MatrixVectorHomogen(grid, L.CellCoefficient(a, c)).multiplication(x)
        fa = lambda p: float(gen_coef(p[None, :])[0][0])
        fc = lambda p: float(gen_coef(p[None, :])[1][0])
        y_fun = L.MatrixVectorHomogen(grid, L.HostFunction(fa, fc)).multiplication(x)
        assert np.array_equal(y_cell, y_fun)
        n = a.size
        one = np.ones(n)
        for st_cell, st_ref in ((L.CellCoefficient(a=one), L.Poisson()), (L.CellCoefficient(c=one), L.HelmHoltz())):
            yc = L.MatrixVectorHomogen(grid, st_cell).multiplication(x)
            yr = L.MatrixVectorHomogen(grid, st_ref).multiplication(x)
            assert relmax(yc, yr) < 1e-12


def test_cell_coefficient_device_layout_regular_grid():
    """On a regular grid the cell-wise operator keeps the grid's device order: device
    tensors in lsg_grid_device_order give the host-slot results; a batch of two
    instances with different cell values matches single instances bit for bit."""
    import torch
    grid = L.AdaptiveSparseGrid.regular(4, 3)
    a, c = cell_values(grid)
    B = 2
    op = L.MatrixVectorHomogen(grid, L.CellCoefficient(a, c), batch=B)
    op.set_coefficient(1, L.CellCoefficient(2.0 * a, None if c is None else 0.5 * c))
    rng = np.random.default_rng(9)
    xh = rng.standard_normal((B, grid.size))
    dev = grid.device_order()
    xd = np.zeros_like(xh)
    xd[:, dev] = xh
    x = torch.tensor(xd, device="cuda")
    y = torch.zeros_like(x)
    op.multiplication(x, y)
    yd = y.cpu().numpy()
    for b, (ab, cb) in enumerate(((a, c), (2.0 * a, 0.5 * c))):
        single = L.MatrixVectorHomogen(grid, L.CellCoefficient(ab, cb))
        ys = single.multiplication(xh[b])
        assert np.array_equal(ys, yd[b][dev])
        o = Oracle(*regular_grid(4, 3))
        # the oracle's gen pair scaled per instance: compare the first instance only
        if b == 0:
            assert relmax(ys, o.set_coef(kind=COEF_GEN).matvec(xh[b])) < 1e-12


def test_cell_coefficient_rejects_bad_input():
    grid = L.AdaptiveSparseGrid.regular(2, 3)
    n = grid.cells()[0].shape[0]
    with pytest.raises(L.LoadsgError):
        L.MatrixVectorHomogen(grid, L.CellCoefficient(np.ones(n - 1)))
    bad = np.ones(n)
    bad[3] = np.nan
    with pytest.raises(L.LoadsgError):
        L.MatrixVectorHomogen(grid, L.CellCoefficient(bad))


def test_sparse_graph_sees_new_coefficients():
    """After set_coefficient, a solve that replays the captured CG chunk uses the new
    coefficients: mixed-precision solves on an adaptive grid, coefficients changed
    between two solves, agree with a fresh operator pair (iterations equal, x bit for
    bit)."""
    import torch
    gv = load("d3_adaptive_expsin")
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    rng = np.random.default_rng(12)
    xi1, xi2 = rng.normal(0.0, 0.5, 6), rng.normal(0.0, 0.5, 6)
    st1 = L.VariableCoefficient(L.COEF_EXPSIN, xi=xi1)
    st2 = L.VariableCoefficient(L.COEF_EXPSIN, xi=xi2)
    op64, op32 = L.MatrixVectorHomogen(grid, st1), L.MatrixVectorHomogen(grid, st1, precision=32)
    f = torch.zeros(1, grid.size, dtype=torch.float64, device="cuda")
    op64.rhs_sines(np.array([[1.0, -0.5]]), np.array([1, 2], dtype=np.int32), out=f)
    x1 = torch.zeros_like(f)
    L.CG.solveMixed(1e-10, x1, f, op64, op32)
    op64.set_coefficient(0, st2)
    op32.set_coefficient(0, st2)
    x2 = torch.zeros_like(f)
    info2 = L.CG.solveMixed(1e-10, x2, f, op64, op32)
    fresh64, fresh32 = L.MatrixVectorHomogen(grid, st2), L.MatrixVectorHomogen(grid, st2, precision=32)
    x3 = torch.zeros_like(f)
    info3 = L.CG.solveMixed(1e-10, x3, f, fresh64, fresh32)
    assert info2["iterations"] == info3["iterations"] and info2["converged"] == 1
    assert torch.equal(x2, x3)
    # and the same for the FP64 CG with the reference stop rule on a cell coefficient
    a, c = cell_values(grid)
    opc = L.MatrixVectorHomogen(grid, L.CellCoefficient(a, c))
    fh = gv["f_rhs"]
    xa = np.zeros(grid.size)
    L.CG.solveHomogen(1e-10, xa, fh, opc)
    opc.set_coefficient(0, L.CellCoefficient(3.0 * a, c))
    xb = np.zeros(grid.size)
    ib = L.CG.solveHomogen(1e-10, xb, fh, opc)
    xc = np.zeros(grid.size)
    ic = L.CG.solveHomogen(1e-10, xc, fh, L.MatrixVectorHomogen(grid, L.CellCoefficient(3.0 * a, c)))
    assert ib["iterations"] == ic["iterations"]
    assert np.array_equal(xb, xc)
