"""Error statuses of the C ABI on the GPU path: a NaN right-hand side (LSG_ERR_NAN),
CG stopped at maxit (LSG_ERR_NOCONV, with the reference's iteration counting:
cg_method.h:852 runs at most maxIteration iterations), refinement sweeps exhausted, and
host arrays of the wrong batch shape.  Every case is a bounded host-side condition;
no kernel faults or hangs."""
import ctypes

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import Oracle, regular_grid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


def _problem(d=3, level=4):
    grid = L.AdaptiveSparseGrid.regular(d, level)
    op = L.MatrixVectorHomogen(grid, L.Poisson())
    f = op.rhs_sines(np.array([[1.0, 0.5]]), np.array([1, 2], dtype=np.int32))
    return grid, op, f


def test_nan_rhs_returns_nan_status():
    grid, op, f = _problem()
    f = f.copy()
    f[5] = np.nan
    with pytest.raises(L.LoadsgError, match="error -5"):
        L.CG.solveHomogen(1e-10, np.zeros(grid.size), f, op)


def test_maxit_returns_noconv_with_reference_counting():
    """The raw status is LSG_ERR_NOCONV; the solve stops after exactly maxit iterations,
    and its iterate equals the oracle's after the same number of iterations."""
    grid, op, f = _problem()
    info = (L.SolveInfo * 1)()
    x = np.zeros(grid.size)
    rc = L.lib().lsg_cg_solve(op._h, L._ptr(f), L._ptr(x), 1e-10, L.STOP_REF_ABS, 3, info, L.LSG_HOST_SLOT)
    assert rc == L.LSG_ERR_NOCONV
    assert info[0].iterations == 3 and info[0].converged == 0
    xo, ito, _ = Oracle(*regular_grid(3, 4)).cg(f, eps=1e-10, maxit=3)
    assert ito == 3
    assert float(np.max(np.abs(x - xo)) / np.max(np.abs(xo))) < 1e-12
    # the Python mirror returns normally (the reference returns true after maxIteration)
    x2 = np.zeros(grid.size)
    res = L.CG.solveHomogen(1e-10, x2, f, op, maxit=3)
    assert res["converged"] == 0 and res["iterations"] == 3
    # enough iterations: converged, status OK
    rc = L.lib().lsg_cg_solve(op._h, L._ptr(f), L._ptr(np.zeros(grid.size)), 1e-10, L.STOP_REF_ABS, 500, info,
                              L.LSG_HOST_SLOT)
    assert rc == 0 and info[0].converged == 1


def test_refinement_sweeps_exhausted_returns_noconv():
    grid = L.AdaptiveSparseGrid.regular(3, 4)
    st = L.VariableCoefficient(L.COEF_CFG3)
    op64, op32 = L.MatrixVectorHomogen(grid, st), L.MatrixVectorHomogen(grid, st, precision=32)
    f = op64.rhs_sines(np.array([[1.0]]), np.array([1], dtype=np.int32))
    info = (L.SolveInfo * 1)()
    x = np.zeros(grid.size)
    rc = L.lib().lsg_ir_solve(op64._h, op32._h, L._ptr(f), L._ptr(x), 1e-12, 1e-1, 1, info, L.LSG_HOST_SLOT)
    assert rc == L.LSG_ERR_NOCONV and info[0].converged == 0
    res = L.CG.solveMixed(1e-12, np.zeros(grid.size), f, op64, op32, inner_rel_tol=1e-1, maxit=1)
    assert res["converged"] == 0


def test_host_arrays_of_the_wrong_batch_shape_are_rejected():
    grid = L.AdaptiveSparseGrid.regular(2, 4)
    op = L.MatrixVectorHomogen(grid, L.Poisson(), batch=2)
    with pytest.raises(L.LoadsgError, match="expected shape"):
        op.multiplication(np.zeros(grid.size))                  # one instance for a batch of two
    y = op.multiplication(np.ones((2, grid.size)))
    assert y.shape == (2, grid.size) and np.array_equal(y[0], y[1])
    import torch
    with pytest.raises(L.LoadsgError, match="GPU tensor"):
        op.multiplication(torch.zeros(2, grid.size, dtype=torch.float32, device="cuda"),
                          torch.zeros(2, grid.size, dtype=torch.float32, device="cuda"))
