"""CG iteration chunks are captured once as a hipGraph and replayed.  Replays (many per
solve, several solves in a row, with eager matvecs between them) must give exactly the
results of direct launches (LSG_NO_GRAPH=1): same iterations, bit-identical solutions."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solve_all(L, torch, grid, xis, f, eager_x):
    B = f.shape[0]
    op64 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=32, batch=B)
    for b in range(B):
        c = L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b])
        op64.set_coefficient(b, c)
        op32.set_coefficient(b, c)
    out = []
    for rep in range(3):
        x = torch.zeros_like(f)
        info = L.CG.solveHomogen(1e-10, x, f, op64, stop_rule=L.STOP_REL, maxit=2000)
        y = torch.zeros_like(eager_x)
        op64.multiplication(eager_x, y)            # eager work between the replays
        torch.cuda.synchronize()
        xm = torch.zeros_like(f)
        im = L.CG.solveMixed(1e-9, xm, f, op64, op32)
        out.append((x.clone(), [i["iterations"] for i in info], xm.clone(), [i["iterations"] for i in im], y.clone()))
    return out


def test_graph_replays_match_direct_launches():
    import torch
    import loadsg_amd as L
    print("graph replays d=4 L5 B=5", flush=True)
    rng = np.random.default_rng(11)
    B = 5
    grid = L.AdaptiveSparseGrid.regular(4, 5)
    xis = rng.normal(0.0, 0.5, (B, 8))
    f = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    ex = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    os.environ.pop("LSG_NO_GRAPH", None)
    with_graph = _solve_all(L, torch, grid, xis, f, ex)
    os.environ["LSG_NO_GRAPH"] = "1"
    try:
        direct = _solve_all(L, torch, grid, xis, f, ex)
    finally:
        os.environ.pop("LSG_NO_GRAPH", None)
    ref = direct[0]
    for got in with_graph + direct:
        assert got[1] == ref[1] and got[3] == ref[3]
        assert torch.equal(got[0], ref[0])
        assert torch.equal(got[2], ref[2])
        assert torch.equal(got[4], ref[4])
