"""Warm-started mixed-precision solves: a learned start converges to the same
tolerance, and a start worse than zero is replaced by zero (bit-identical to the
zero-start solve)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(B, d=3, level=5, nxi=4, seed=3):
    import torch
    import loadsg_amd as L
    rng = np.random.default_rng(seed)
    xi = rng.normal(0.0, 0.5, (B, nxi))
    eta = rng.standard_normal((B, 2))
    m = np.array([1, 2], dtype=np.int32)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    op64 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[0]), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[0]), precision=32, batch=B)
    for b in range(B):
        s = L.VariableCoefficient(L.COEF_EXPSIN, xi=xi[b])
        op64.set_coefficient(b, s)
        op32.set_coefficient(b, s)
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    op64.rhs_sines(eta, m, out=f)
    return L, grid, xi, eta, m, op64, op32, f


def test_bad_start_falls_back_to_zero():
    import torch
    print("fallback d=3 L5 B=6", flush=True)
    L, grid, xi, eta, m, op64, op32, f = _setup(6)
    x0 = torch.zeros_like(f)
    i0 = L.CG.solveMixed(1e-8, x0, f, op64, op32)
    xb = torch.randn_like(f) * 1e6
    xb[3].zero_()                       # one instance starts from zero anyway
    ib = L.CG.solveMixed(1e-8, xb, f, op64, op32)
    assert torch.equal(xb, x0)
    assert [i["iterations"] for i in ib] == [i["iterations"] for i in i0]
    assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-8 for i in ib)


def test_learned_start_converges():
    import torch
    from loadsg_amd.warmstart import WarmStart
    print("warm start d=3 L5", flush=True)
    L, grid, xi, eta, m, op64, op32, f = _setup(8)
    ws = WarmStart(grid, m, n_xi=4, width=128, hidden=2, seed=1)
    tr = np.random.default_rng(4).normal(0.0, 0.5, (128, 4))
    U = ws.basis_solutions(tr, batch=64)
    assert U.shape == (128, 2, grid.size)
    losses = ws.fit(tr, U, steps=1500, batch=64, lr=3e-3)
    assert np.mean(losses[-20:]) < 0.5 * losses[0]
    x = ws.predict(xi, eta)
    assert x.dtype == torch.float64 and x.shape == f.shape
    Ax = torch.zeros_like(f)
    op64.multiplication(x, Ax)
    rel0 = (torch.linalg.norm(f - Ax, dim=1) / torch.linalg.norm(f, dim=1)).cpu().numpy()
    info = L.CG.solveMixed(1e-8, x, f, op64, op32)
    assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-8 for i in info)
    xr = torch.zeros_like(f)
    L.CG.solveHomogen(1e-12, xr, f, op64, stop_rule=L.STOP_REL, maxit=2000)
    assert float(torch.max(torch.linalg.norm(x - xr, dim=1) / torch.linalg.norm(xr, dim=1))) < 1e-7
    # what the network gains over the zero start on these held-out coefficients
    x0 = torch.zeros_like(f)
    info0 = L.CG.solveMixed(1e-8, x0, f, op64, op32)
    warm_it = sum(i["iterations"] for i in info)
    zero_it = sum(i["iterations"] for i in info0)
    print("start residuals", rel0, "iterations warm %d zero %d" % (warm_it, zero_it), flush=True)
    assert np.median(rel0) < 0.9          # the start is better than zero (relative residual 1)
    assert warm_it <= zero_it
