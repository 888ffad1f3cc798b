"""Batches of 16 or more take the fine apply size classes and, inside CG, grids compacted
to the instances still iterating.  Neither may change results: a batch of 20 must match
20 single-instance operators (coarse classes, no compaction) bit for bit, for the
separable kernel (exp-sine coefficient) and the multi-term one (CFG2), in the matvec and
in the mixed-precision solve (where instances finish after different iteration counts)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["expsin", "cfg2"])
def test_fine_classes_and_compaction_match_single(kind):
    import torch
    import loadsg_amd as L
    print("batch 20 d=4 L5", kind, flush=True)
    rng = np.random.default_rng(21)
    B = 20
    grid = L.AdaptiveSparseGrid.regular(4, 5)
    xis = rng.normal(0.0, 0.5, (B, 8))

    def coef(b):
        return L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b]) if kind == "expsin" else \
            L.VariableCoefficient(L.COEF_CFG2)

    op64 = L.MatrixVectorHomogen(grid, coef(0), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, coef(0), precision=32, batch=B)
    for b in range(B):
        op64.set_coefficient(b, coef(b))
        op32.set_coefficient(b, coef(b))
    x = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    y = torch.zeros_like(x)
    op64.multiplication(x, y)
    # right-hand sides of very different size: instances converge after different counts
    f = torch.tensor(rng.standard_normal((B, grid.size)) * np.logspace(-3, 3, B)[:, None], device="cuda")
    xm = torch.zeros_like(f)
    info = L.CG.solveMixed(1e-9, xm, f, op64, op32)
    its = [i["iterations"] for i in info]
    for b in range(B):
        s64 = L.MatrixVectorHomogen(grid, coef(b), precision=64, batch=1)
        s32 = L.MatrixVectorHomogen(grid, coef(b), precision=32, batch=1)
        ys = torch.zeros_like(x[b:b + 1])
        s64.multiplication(x[b:b + 1].contiguous(), ys)
        assert torch.equal(ys[0], y[b])
        xs = torch.zeros_like(f[b:b + 1])
        i1 = L.CG.solveMixed(1e-9, xs, f[b:b + 1].contiguous(), s64, s32)
        assert i1["iterations"] == its[b]
        assert torch.equal(xs[0], xm[b])


def test_refinement_subsets_match_single():
    """Starts of different quality make instances leave the refinement after different
    sweeps (an exact start after none, the others after one or more; a start worse than
    zero is replaced by zero); from the second sweep on, the FP64 residuals and the inner
    solves are then run over the unconverged instances only.  Each instance must match
    its single-instance solve bit for bit: x, sweeps, inner iterations and the residual."""
    import torch
    import loadsg_amd as L
    print("batch 20 d=4 L5 mixed starts", flush=True)
    rng = np.random.default_rng(22)
    B = 20
    grid = L.AdaptiveSparseGrid.regular(4, 5)
    xis = rng.normal(0.0, 0.5, (B, 8))

    def coef(b):
        return L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b])

    op64 = L.MatrixVectorHomogen(grid, coef(0), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, coef(0), precision=32, batch=B)
    for b in range(B):
        op64.set_coefficient(b, coef(b))
        op32.set_coefficient(b, coef(b))
    f = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    xr = torch.zeros_like(f)
    L.CG.solveHomogen(1e-13, xr, f, op64, stop_rule=L.STOP_REL, maxit=5000)
    x0 = torch.zeros_like(f)
    kind = [b % 4 for b in range(B)]
    for b in range(B):
        if kind[b] == 0:
            x0[b] = xr[b]
        elif kind[b] == 1:
            x0[b] = xr[b] * (1.0 + 1e-5 * torch.tensor(rng.standard_normal(grid.size), device="cuda"))
        elif kind[b] == 3:
            x0[b] = -3.0 * xr[b]
    xm = x0.clone()
    info = L.CG.solveMixed(1e-9, xm, f, op64, op32)
    outer = [i["outer"] for i in info]
    assert {outer[b] for b in range(B) if kind[b] == 0} == {0}
    assert max(outer) >= 2, outer
    for b in range(B):
        s64 = L.MatrixVectorHomogen(grid, coef(b), precision=64, batch=1)
        s32 = L.MatrixVectorHomogen(grid, coef(b), precision=32, batch=1)
        xs = x0[b:b + 1].clone()
        i1 = L.CG.solveMixed(1e-9, xs, f[b:b + 1].contiguous(), s64, s32)
        assert (i1["outer"], i1["iterations"]) == (outer[b], info[b]["iterations"]), b
        assert i1["rel_residual"] == info[b]["rel_residual"], b
        assert info[b]["converged"] == 1 and info[b]["rel_residual"] <= 1e-9
        assert torch.equal(xs[0], xm[b]), b


def test_set_coefficients_equals_one_by_one():
    """lsg_op_set_coefs (tables built on host threads, one upload) gives the operator that
    lsg_op_set_coef instance by instance gives, bit for bit, in FP64 and FP32."""
    import torch
    import loadsg_amd as L
    grid = L.AdaptiveSparseGrid.regular(4, 5)
    B = 7
    rng = np.random.default_rng(11)
    cs = [L.VariableCoefficient(L.COEF_EXPSIN, xi=rng.normal(0.0, 0.5, 16)) for _ in range(B)]
    for prec, dt in ((64, torch.float64), (32, torch.float32)):
        a = L.MatrixVectorHomogen(grid, cs[0], precision=prec, batch=B)
        b = L.MatrixVectorHomogen(grid, cs[0], precision=prec, batch=B)
        a.set_coefficients(cs[1:], first=1)
        for i in range(1, B):
            b.set_coefficient(i, cs[i])
        x = torch.randn(B, grid.size, dtype=dt, device="cuda")
        ya, yb = torch.zeros_like(x), torch.zeros_like(x)
        a.multiplication(x, ya)
        b.multiplication(x, yb)
        assert torch.equal(ya, yb)
