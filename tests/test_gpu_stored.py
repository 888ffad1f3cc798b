"""Stored element matrices on the GPU (LSG_COEF_MATRIX / LSG_COEF_INTEGRATOR): the
counterpart of LocalStiffnessMatrices (localStiffnessMatrices/LocalStiffnessMatrices.h:76-321)
and of MatrixVectorHomogen::multiplication's StoreLocalStiffnessMatrix branch
(MatrixVectorHomogen.h:170-245,1619-1700).

The element matrices are the harness's quadrature plug-in (QuadStencil, a and c varying
inside each cell), computed here for the library's cells.  Checked against the compiled
reference's golden vectors (LocalStiffnessMatricesMemoryDistribution built from the same
plug-in; the diagonal of precondLocalStiffnessMatrix) on regular (d = 3, 5: the cell-wise
and the row-gather kernels) and adaptive grids, and against the CPU oracle with the same
matrices.  Tolerances: FP64 1e-12 relative (max-norm), FP32 2e-5, CG iteration counts exact,
CG solutions 1e-10.
"""
import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import Oracle, gen_coef, quad_element_matrices
from test_gpu_parity import load, relmax

pytestmark = pytest.mark.gpu

CASES = ["d3_l4_gen", "d5_l3_gen", "d3_adaptive_gen"]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


def stored(grid, scale=1.0):
    lev, pos = grid.cell_boxes()
    coef = lambda x: tuple(scale * v for v in gen_coef(x))
    return L.LocalStiffnessMatrices(grid, quad_element_matrices(lev, pos, coef))


def oracle_for(gv, scale=1.0):
    o = Oracle(gv["idx"], gv["active"])
    lev, pos = o.cell_boxes()
    coef = lambda x: tuple(scale * v for v in gen_coef(x))
    return o.set_cell_matrices(quad_element_matrices(lev, pos, coef))


@pytest.mark.parametrize("case", CASES)
def test_stored_matvec_and_diagonal_golden(case):
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    op = L.MatrixVectorHomogen(grid, stored(grid))
    y = op.multiplication(gv["x"])
    assert relmax(y, gv["Ax_stored"]) < 1e-12
    act = gv["active"] == 1
    assert np.all(y[~act] == 0.0)
    z = op.diagonal()
    assert relmax(z[act], gv["diag_stored"][act]) < 1e-12
    assert np.all(z[~act] == 1.0)
    op32 = L.MatrixVectorHomogen(grid, stored(grid), precision=32)
    assert relmax(op32.multiplication(gv["x"]), gv["Ax_stored"]) < 2e-5


def test_integrator_callback_equals_stored_entries():
    """LSG_COEF_INTEGRATOR (the library calls the host integrator per cell and interior
    pair) builds the same operator as the entries passed in bulk."""
    gv = load("d3_l4_gen")
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])

    def integ(level, pos, p, q):
        E = quad_element_matrices(level[None, :], pos[None, :])
        return E[0, L.cell_pairs(len(level)).index((p, q))]

    y1 = L.MatrixVectorHomogen(grid, L.CellIntegrator(integ)).multiplication(gv["x"])
    y2 = L.MatrixVectorHomogen(grid, stored(grid)).multiplication(gv["x"])
    assert np.array_equal(y1, y2)


def test_centre_matrices_equal_cell_coefficient_operator():
    """Stored a(centre) K + c(centre) M equals the LSG_COEF_CELL operator of the same a, c."""
    from test_oracle_stored import centre_element_matrices
    gv = load("d3_adaptive_gen")
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    lev, pos = grid.cell_boxes()
    x, _ = grid.cells()
    a, c = gen_coef(x)
    y1 = L.MatrixVectorHomogen(grid, L.LocalStiffnessMatrices(grid, centre_element_matrices(lev, pos))).multiplication(gv["x"])
    y2 = L.MatrixVectorHomogen(grid, L.CellCoefficient(a, c)).multiplication(gv["x"])
    assert relmax(y1, y2) < 1e-13
    assert relmax(y1, gv["Ax_var"]) < 1e-12


@pytest.mark.parametrize("case", ["d3_l4_gen", "d3_adaptive_gen"])
def test_stored_cg_vs_oracle(case):
    """Reference stopping rule r^T D^-1 r < eps^2 with the stored-matrix diagonal:
    iteration counts equal to the oracle's, solutions within 1e-10."""
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    op = L.MatrixVectorHomogen(grid, stored(grid))
    x = np.zeros(grid.size)
    info = L.CG.solveHomogen(1e-10, x, gv["f_rhs"], op, stop_rule=L.STOP_REF_ABS, maxit=3000)
    xo, ito, _ = oracle_for(gv).cg(gv["f_rhs"], eps=1e-10, maxit=3000)
    assert info["iterations"] == ito
    assert relmax(x, xo) < 1e-10


def test_stored_batch_instances_and_recoefficient():
    """Three instances with different matrices match single-instance operators bit for bit;
    new matrices on an operator that has already solved (replayed CG graphs) give the
    solution of a fresh operator; the mixed-precision solve reaches 1e-10 and lies within
    1e-8 (relative 2-norm) of a tight FP64 solve."""
    gv = load("d3_adaptive_gen")
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    scales = [1.0, 0.5, 2.5]
    B = len(scales)
    op = L.MatrixVectorHomogen(grid, stored(grid, scales[0]), batch=B)
    for b in range(1, B):
        op.set_coefficient(b, stored(grid, scales[b]))
    X = np.tile(gv["x"], (B, 1))
    Y = op.multiplication(X)
    for b in range(B):
        yb = L.MatrixVectorHomogen(grid, stored(grid, scales[b])).multiplication(gv["x"])
        assert np.array_equal(Y[b], yb)
    # solve, then change instance 0's matrices and solve again
    F = np.tile(gv["f_rhs"], (B, 1))
    X0 = np.zeros((B, grid.size))
    L.CG.solveHomogen(1e-10, X0, F, op, stop_rule=L.STOP_REF_ABS, maxit=3000)
    op.set_coefficient(0, stored(grid, 3.0))
    X1 = np.zeros((B, grid.size))
    info = L.CG.solveHomogen(1e-10, X1, F, op, stop_rule=L.STOP_REF_ABS, maxit=3000)
    fresh = L.MatrixVectorHomogen(grid, stored(grid, 3.0))
    xf = np.zeros(grid.size)
    inf = L.CG.solveHomogen(1e-10, xf, gv["f_rhs"], fresh, stop_rule=L.STOP_REF_ABS, maxit=3000)
    assert info[0]["iterations"] == inf["iterations"]
    assert relmax(X1[0], xf) < 1e-10
    # mixed precision: FP32 inner PCG on the stored matrices under FP64 refinement
    op32 = L.MatrixVectorHomogen(grid, stored(grid, 3.0), precision=32)
    xm = np.zeros(grid.size)
    im = L.CG.solveMixed(1e-10, xm, gv["f_rhs"], fresh, op32)
    assert im["converged"] == 1 and im["rel_residual"] <= 1e-10
    xt = np.zeros(grid.size)
    L.CG.solveHomogen(1e-13, xt, gv["f_rhs"], fresh, stop_rule=L.STOP_REL, maxit=5000)
    assert np.linalg.norm(xm - xt) / np.linalg.norm(xt) < 1e-8
