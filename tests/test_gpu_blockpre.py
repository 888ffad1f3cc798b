"""The depth-block preconditioner on the GPU (lsg_op_set_precond, LSG_PRECOND_DEPTH_BLOCK):
no counterpart in the reference, checked against its numpy restatement
(tests/blockpre_ref.py, pinned on the CPU by tests/test_blockpre_ref.py) driven by the same
operator and diagonal.  PCG to a relative residual of 1e-10: iteration counts within one of the
restatement's, solutions within 1e-9.  Both ways of applying a block are exercised: dense
inverses (up to 64 points on FP64 operators) and the fast diagonalisation of larger blocks
(d = 2 level 8 and 9, d = 3 level 8: blocks of 128 and 256 points); the FP32 operator of the
mixed solve keeps dense blocks of up to 128 points (d = 5, level 7).  The mixed-precision solve
with the preconditioner on its FP32 operator reaches 1e-10 in fewer inner iterations than with
the diagonal; batches match single instances bit for bit; graph replays match direct launches
bit for bit.  On adaptive grids (hanging points, partly filled depth blocks) the blocks are
taken over the active points, and the fast diagonalisation over product sets of 1D positions.
"""
import os

import numpy as np
import pytest

import loadsg_amd as L
from blockpre_ref import make_preconditioner, pcg
from test_gpu_parity import relmax

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


def expsin(seed):
    return L.VariableCoefficient(L.COEF_EXPSIN, xi=np.random.default_rng(seed).normal(0.0, 0.5, 16))


@pytest.mark.parametrize("d,level", [(2, 8), (2, 9), (3, 5), (3, 8), (5, 4)])
def test_block_pcg_matches_restatement(d, level):
    print("d=%d level=%d" % (d, level), flush=True)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    idx, _ = grid.slot_order()
    op = L.MatrixVectorHomogen(grid, expsin(d + level))
    D = op.diagonal()
    b = np.random.default_rng(level).standard_normal(grid.size)
    xr, itr = pcg(op.multiplication, b, make_preconditioner(idx, D, max_dense=64), 1e-10)
    op.set_precond(L.PRECOND_DEPTH_BLOCK)
    x = np.zeros(grid.size)
    info = L.CG.solveHomogen(1e-10, x, b, op, stop_rule=L.STOP_REL, maxit=2000)
    assert info["converged"] == 1
    assert abs(info["iterations"] - itr) <= 1, (info["iterations"], itr)
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
    op.set_precond(L.PRECOND_DIAG)
    xd = np.zeros(grid.size)
    infod = L.CG.solveHomogen(1e-10, xd, b, op, stop_rule=L.STOP_REL, maxit=2000)
    assert info["iterations"] < infod["iterations"]
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


@pytest.mark.parametrize("level", [5, 7])
def test_mixed_solve_with_block_preconditioner(level):
    """d = 5: level 5 (blocks up to 32 points), level 7 (blocks of 128 points on the FP32
    operator)."""
    import torch
    print("d=5 level=%d" % level, flush=True)
    grid = L.AdaptiveSparseGrid.regular(5, level)
    B = 4
    ops = {}
    for prec in (64, 32):
        ops[prec] = L.MatrixVectorHomogen(grid, expsin(0), precision=prec, batch=B)
        for b in range(B):
            ops[prec].set_coefficient(b, expsin(b))
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    ops[64].rhs_sines(np.random.default_rng(3).standard_normal((B, 8)),
                      np.array([1, 2, 3, 1, 2, 3, 1, 2], dtype=np.int32), out=f)
    res = {}
    for kind in (L.PRECOND_DIAG, L.PRECOND_DEPTH_BLOCK):
        ops[32].set_precond(kind)
        x = torch.zeros_like(f)
        info = L.CG.solveMixed(1e-10, x, f, ops[64], ops[32])
        assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-10 for i in info)
        res[kind] = (x, np.mean([i["iterations"] for i in info]))
    print("inner iterations diag %.1f, depth-block %.1f" % (res[0][1], res[1][1]), flush=True)
    assert res[L.PRECOND_DEPTH_BLOCK][1] < 0.6 * res[L.PRECOND_DIAG][1]
    xd, xb = res[0][0], res[1][0]
    assert float(torch.max(torch.linalg.norm(xd - xb, dim=1) / torch.linalg.norm(xd, dim=1))) < 1e-8


def _adaptive_grid(name):
    if name == "golden":
        from test_gpu_parity import load
        gv = load("d3_adaptive_cfg3")
        return np.asarray(gv["idx"]), np.asarray(gv["active"])
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "grids", name + ".npz"))
    return z["idx"], z["active"]


@pytest.mark.parametrize("name", ["golden", "cfg3_start4_level8", "cfg3_start4_level10"])
def test_block_pcg_on_adaptive_grids(name):
    """Grids refined towards a corner (hanging points, partly filled depth blocks): dense
    inverses of the small blocks (hanging points as the identity), fast diagonalisation of
    the product sets of the larger ones in a point order that differs from the device order;
    the cfg3 coefficient 1 + |x|^2."""
    print("adaptive", name, flush=True)
    idx, act = _adaptive_grid(name)
    grid = L.AdaptiveSparseGrid(idx, act)
    op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_CFG3))
    D = op.diagonal()
    assert np.all(D[act == 0] == 1.0)
    b = np.random.default_rng(len(idx)).standard_normal(len(idx))
    b[act == 0] = 0.0
    xr, itr = pcg(op.multiplication, b, make_preconditioner(idx, D, max_dense=64, active=act), 1e-10)
    op.set_precond(L.PRECOND_DEPTH_BLOCK)
    x = np.zeros(len(idx))
    info = L.CG.solveHomogen(1e-10, x, b, op, stop_rule=L.STOP_REL, maxit=3000)
    assert info["converged"] == 1
    assert abs(info["iterations"] - itr) <= 1, (info["iterations"], itr)
    assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)
    assert np.all(x[act == 0] == 0.0)
    op.set_precond(L.PRECOND_DIAG)
    xd = np.zeros(len(idx))
    infod = L.CG.solveHomogen(1e-10, xd, b, op, stop_rule=L.STOP_REL, maxit=3000)
    print("iterations diag %d, depth-block %d" % (infod["iterations"], info["iterations"]), flush=True)
    assert info["iterations"] < 0.7 * infod["iterations"]
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


def test_mixed_solve_with_block_preconditioner_on_an_adaptive_grid():
    import torch
    idx, act = _adaptive_grid("cfg3_start4_level8")
    grid = L.AdaptiveSparseGrid(idx, act)
    B = 4
    ops = {p: L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_CFG3), precision=p, batch=B)
           for p in (64, 32)}
    # device tensors hold the device order: hanging points at grid.device_order()[slot]
    hang = np.zeros(len(idx), bool)
    hang[grid.device_order()] = act == 0
    f = torch.tensor(np.random.default_rng(9).standard_normal((B, len(idx))), device="cuda")
    f[:, torch.tensor(hang, device="cuda")] = 0.0
    res = {}
    for kind in (L.PRECOND_DIAG, L.PRECOND_DEPTH_BLOCK):
        ops[32].set_precond(kind)
        x = torch.zeros_like(f)
        info = L.CG.solveMixed(1e-10, x, f, ops[64], ops[32])
        assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-10 for i in info)
        res[kind] = (x, np.mean([i["iterations"] for i in info]))
    print("inner iterations diag %.1f, depth-block %.1f" % (res[0][1], res[1][1]), flush=True)
    assert res[L.PRECOND_DEPTH_BLOCK][1] < 0.6 * res[L.PRECOND_DIAG][1]
    xd, xb = res[0][0], res[1][0]
    assert float(torch.max(torch.linalg.norm(xd - xb, dim=1) / torch.linalg.norm(xd, dim=1))) < 1e-8


@pytest.mark.parametrize("kind", ["regular", "adaptive"])
def test_block_batch_matches_single_and_graph_replay(kind):
    import torch
    if kind == "regular":
        grid = L.AdaptiveSparseGrid.regular(3, 6)
        act = np.ones(grid.size, np.uint8)
    else:
        idx, act = _adaptive_grid("golden")
        grid = L.AdaptiveSparseGrid(idx, act)
    B = 3
    op = L.MatrixVectorHomogen(grid, expsin(0), batch=B).set_precond(L.PRECOND_DEPTH_BLOCK)
    for b in range(B):
        op.set_coefficient(b, expsin(10 + b))
    F = np.random.default_rng(5).standard_normal((B, grid.size))
    F[:, act == 0] = 0.0
    X = np.zeros((B, grid.size))
    infos = L.CG.solveHomogen(1e-10, X, F, op, stop_rule=L.STOP_REL, maxit=2000)
    for b in range(B):
        single = L.MatrixVectorHomogen(grid, expsin(10 + b)).set_precond(L.PRECOND_DEPTH_BLOCK)
        xs = np.zeros(grid.size)
        info = L.CG.solveHomogen(1e-10, xs, F[b], single, stop_rule=L.STOP_REL, maxit=2000)
        assert info["iterations"] == infos[b]["iterations"]
        assert np.array_equal(xs, X[b])
    # direct launches instead of the captured chunk graphs give the same bits
    os.environ["LSG_NO_GRAPH"] = "1"
    try:
        X2 = np.zeros((B, grid.size))
        L.CG.solveHomogen(1e-10, X2, F, op, stop_rule=L.STOP_REL, maxit=2000)
    finally:
        del os.environ["LSG_NO_GRAPH"]
    assert np.array_equal(X, X2)
    # a second graph-replayed solve after the eager one
    X3 = np.zeros((B, grid.size))
    L.CG.solveHomogen(1e-10, X3, F, op, stop_rule=L.STOP_REL, maxit=2000)
    assert np.array_equal(X, X3)


@pytest.mark.parametrize("d,level", [(3, 8), (5, 6), (2, 7)])
def test_fp32_block_application_matches_restatement(d, level):
    """FP32 operators apply their dense blocks of 32 points and more on the matrix cores
    (blockmm_kernel), the smaller ones in blockpre_kernel.  One PCG step from zero gives
    x1 = alpha M^-1 f; its direction must be the restatement's M^-1 f (FP64 inverses) to FP32
    accuracy.  d=3 L8: blocks of 32, 64 and 128 points; d=5 L6: up to 32; d=2 L7: up to 32."""
    print("fp32 block application d=%d level=%d" % (d, level), flush=True)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    idx, _ = grid.slot_order()
    op = L.MatrixVectorHomogen(grid, expsin(7), precision=32).set_precond(L.PRECOND_DEPTH_BLOCK)
    D = np.asarray(op.diagonal(), dtype=np.float64)
    b = np.random.default_rng(d * 10 + level).standard_normal(grid.size)
    x = np.zeros(grid.size)
    L.CG.solveHomogen(1e-30, x, b, op, stop_rule=L.STOP_REL, maxit=1)
    href = make_preconditioner(idx, D, max_dense=128)(b)
    u, v = x / np.linalg.norm(x), href / np.linalg.norm(href)
    if u @ v < 0:
        u = -u
    assert np.linalg.norm(u - v) < 2e-5, np.linalg.norm(u - v)
