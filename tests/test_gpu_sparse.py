"""GPU parity of the sparse path: locally adaptive grids with hanging nodes (the
compiled reference's completeDirichletGrid output) through the C ABI, against the
reference's golden vectors and the CPU oracle.

Tolerances as in test_gpu_parity: FP64 1e-12 relative (max-norm), FP32 2e-5,
iteration counts exact.  The diagonal is compared with the oracle (which equals
(A e_p)_p, see test_oracle.py); the reference's own diagonal on adaptive grids
carries stale hanging-node values in 1-4 % of its entries.
"""
import json
import os

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import COEF_CONST, Oracle
from test_gpu_parity import load, relmax, stencil_of as _stencil_of

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
ADAPTIVE = sorted(c for c in CASES if CASES[c]["argv"][0].startswith("adaptive"))


def stencil_of(argv):
    return _stencil_of([None, None, argv[-1]])     # coefficient spec is the last harness argument


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


@pytest.mark.parametrize("case", ADAPTIVE)
def test_sparse_matvec_golden(case):
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    var, _ = stencil_of(CASES[case]["argv"])
    for st, key in ((L.Poisson(), "Ax_poisson"), (L.HelmHoltz(), "Ax_helm"), (var, "Ax_var")):
        op = L.MatrixVectorHomogen(grid, st)
        y = op.multiplication(gv["x"])
        assert relmax(y, gv[key]) < 1e-12, key
        assert np.all(y[gv["active"] == 0] == 0.0)        # hanging nodes get no value
        op32 = L.MatrixVectorHomogen(grid, st, precision=32)
        assert relmax(op32.multiplication(gv["x"]), gv[key]) < 2e-5, key


@pytest.mark.parametrize("case", ADAPTIVE)
def test_sparse_diagonal_vs_oracle(case):
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    var, ovar = stencil_of(CASES[case]["argv"])
    o = Oracle(gv["idx"], gv["active"])
    act = gv["active"] == 1
    for st, okind in ((L.Poisson(), dict(kind=COEF_CONST, a0=1.0, c0=0.0)), (var, ovar)):
        z = L.MatrixVectorHomogen(grid, st).diagonal()
        zo = o.set_coef(**okind).precond()
        assert relmax(z[act], zo[act]) < 1e-12
        assert np.all(z[~act] == 1.0)


@pytest.mark.parametrize("case", ADAPTIVE)
def test_sparse_cg_vs_oracle(case):
    """Reference stopping rule: iteration counts equal to the oracle's (same diagonal),
    solutions within 1e-10; against the reference's golden solution within 1e-8 (its
    diagonal differs in the stale entries)."""
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    var, ovar = stencil_of(CASES[case]["argv"])
    o = Oracle(gv["idx"], gv["active"])
    for st, okind, key in ((L.Poisson(), dict(kind=COEF_CONST, a0=1.0, c0=0.0), "cg_poisson"),
                           (var, ovar, "cg_var")):
        op = L.MatrixVectorHomogen(grid, st)
        x = np.zeros(grid.size)
        info = L.CG.solveHomogen(1e-10, x, gv["f_rhs"], op, stop_rule=L.STOP_REF_ABS, maxit=3000)
        xo, ito, _ = o.set_coef(**okind).cg(gv["f_rhs"], eps=1e-10)
        assert info["iterations"] == ito
        assert relmax(x, xo) < 1e-10
        assert relmax(x, gv[key + "_x"]) < 1e-8


def test_sparse_batch_and_mixed_solve():
    """Batch instances on an adaptive grid match single instances bit for bit; the
    mixed-precision solve reaches 1e-10 and agrees with the FP64 solve."""
    import torch
    gv = load("d3_adaptive_expsin")
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    rng = np.random.default_rng(2)
    B = 3
    xis = rng.normal(0.0, 0.5, (B, 6))
    op64 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=32, batch=B)
    for b in range(B):
        c = L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b])
        op64.set_coefficient(b, c)
        op32.set_coefficient(b, c)
    x = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    y = torch.zeros_like(x)
    op64.multiplication(x, y)
    dev = grid.device_order()
    for b in range(B):
        single = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b]))
        xs = np.zeros(grid.size)
        xs[:] = x[b].cpu().numpy()[dev]
        ys = single.multiplication(xs)
        assert np.array_equal(ys, y[b].cpu().numpy()[dev])
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    op64.rhs_sines(rng.standard_normal((B, 2)), np.array([1, 2], dtype=np.int32), out=f)
    xm = torch.zeros_like(f)
    info = L.CG.solveMixed(1e-10, xm, f, op64, op32)
    assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-10 for i in info)
    xr = torch.zeros_like(f)
    L.CG.solveHomogen(1e-13, xr, f, op64, stop_rule=L.STOP_REL, maxit=5000)
    assert float(torch.max(torch.linalg.norm(xm - xr, dim=1) / torch.linalg.norm(xr, dim=1))) < 1e-8


def test_refined_cfg3_grid_vs_oracle():
    """A config-3 style grid refined towards the origin (10,315 slots, 216 hanging):
    FP64 matvec within 1e-12 and FP32 within 2e-5 of the oracle with a(x) = 1 + |x|^2,
    the diagonal within 1e-12, and the mixed-precision solve to 1e-10 within 1e-8 of
    an FP64 solve."""
    import torch
    print("cfg3 start 4 level 8", flush=True)
    z = np.load(os.path.join(GOLD, "grids", "cfg3_start4_level8.npz"))
    idx, act = z["idx"], z["active"]
    grid = L.AdaptiveSparseGrid(idx, act)
    o = Oracle(idx, act).set_coef(kind=L.COEF_CFG3)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(grid.size) * act
    yo = o.matvec(x)
    st = L.VariableCoefficient(L.COEF_CFG3)
    op64 = L.MatrixVectorHomogen(grid, st)
    assert relmax(op64.multiplication(x), yo) < 1e-12
    assert relmax(L.MatrixVectorHomogen(grid, st, precision=32).multiplication(x), yo) < 2e-5
    a = act == 1
    assert relmax(op64.diagonal()[a], o.precond()[a]) < 1e-12
    B = 4
    o64 = L.MatrixVectorHomogen(grid, st, precision=64, batch=B)
    o32 = L.MatrixVectorHomogen(grid, st, precision=32, batch=B)
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    o64.rhs_sines(rng.standard_normal((B, 2)), np.array([1, 2], dtype=np.int32), out=f)
    xm = torch.zeros_like(f)
    info = L.CG.solveMixed(1e-10, xm, f, o64, o32)
    assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-10 for i in info)
    xr = torch.zeros_like(f)
    L.CG.solveHomogen(1e-13, xr, f, o64, stop_rule=L.STOP_REL, maxit=5000)
    assert float(torch.max(torch.linalg.norm(xm - xr, dim=1) / torch.linalg.norm(xr, dim=1))) < 1e-8


@pytest.mark.parametrize("case,kind", [("d2_adaptive_cfg2", "mass"), ("d3_adaptive_cfg3", "expsin")])
def test_sparse_instance_groups_match_single(case, kind):
    """Batches above 1 take the cell-wise apply over groups of instances (one index-table
    load per group; LSG_SP_GROUP = 4, 2 or 1): batch 9 (two full groups and one of one) in
    FP64 and FP32 must equal 9 single-instance operators bit for bit, in the matvec and in the
    mixed-precision solve, whose instances converge after different counts (the grids are then
    compacted to the instances still iterating, so groups mix instances)."""
    import torch
    print("groups", case, kind, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    B = 9
    rng = np.random.default_rng(31)
    xis = rng.normal(0.0, 0.5, (B, 8))

    def coef(b):
        if kind == "mass":      # constant a and c per instance: stiffness and mass terms
            return L.Stencil(L.COEF_CONST, a0=0.7 + 0.1 * b, c0=0.3 + 0.2 * b)
        return L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b])

    dev = grid.device_order()
    x = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    ops = {}
    for prec, dt in ((64, torch.float64), (32, torch.float32)):
        op = L.MatrixVectorHomogen(grid, coef(0), precision=prec, batch=B)
        for b in range(B):
            op.set_coefficient(b, coef(b))
        ops[prec] = op
        xb = x.to(dt)
        y = torch.zeros_like(xb)
        op.multiplication(xb, y)
        for b in range(B):
            s = L.MatrixVectorHomogen(grid, coef(b), precision=prec)
            ys = torch.zeros_like(xb[b:b + 1])
            s.multiplication(xb[b:b + 1].contiguous(), ys)
            assert torch.equal(ys[0], y[b]), (prec, b)
    ops[64].rhs_sines(rng.standard_normal((B, 2)) * np.logspace(-2, 2, B)[:, None],
                      np.array([1, 2], dtype=np.int32), out=f)
    xm = torch.zeros_like(f)
    info = L.CG.solveMixed(1e-9, xm, f, ops[64], ops[32])
    for b in range(B):
        s64 = L.MatrixVectorHomogen(grid, coef(b), precision=64)
        s32 = L.MatrixVectorHomogen(grid, coef(b), precision=32)
        xs = torch.zeros_like(f[b:b + 1])
        i1 = L.CG.solveMixed(1e-9, xs, f[b:b + 1].contiguous(), s64, s32)
        assert i1["iterations"] == info[b]["iterations"], b
        assert torch.equal(xs[0], xm[b]), b


def test_cfg3_full_level11_grid():
    """The config-3 grid itself (sum level 11: 65,943 slots, 2,940 hanging): FP32 matvec within
    2e-5 of FP64 (the oracle takes minutes per matvec here; the smaller refined grids pin
    both against it), a batch-5 matvec (a group of four and one of one) equal to single
    instances bit for bit, and the mixed-precision solve with the depth-block preconditioner
    on the FP32 operator to 1e-10 within 1e-8 of a tight FP64 solve, in fewer inner
    iterations than with the diagonal."""
    import torch
    print("cfg3 start 4 level 11", flush=True)
    z = np.load(os.path.join(GOLD, "grids", "cfg3_start4_level11.npz"))
    idx, act = z["idx"], z["active"]
    grid = L.AdaptiveSparseGrid(idx, act)
    st = L.VariableCoefficient(L.COEF_CFG3)
    B = 5
    o64 = L.MatrixVectorHomogen(grid, st, precision=64, batch=B)
    o32 = L.MatrixVectorHomogen(grid, st, precision=32, batch=B)
    hang = np.zeros(grid.size, bool)
    hang[grid.device_order()] = act == 0
    rng = np.random.default_rng(11)
    x = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    x[:, torch.tensor(hang, device="cuda")] = 0.0
    y64, y32 = torch.zeros_like(x), torch.zeros(B, grid.size, dtype=torch.float32, device="cuda")
    o64.multiplication(x, y64)
    o32.multiplication(x.float(), y32)
    assert float(torch.max(torch.abs(y32.double() - y64)) / torch.max(torch.abs(y64))) < 2e-5
    s64 = L.MatrixVectorHomogen(grid, st, precision=64)
    for b in (0, 4):
        ys = torch.zeros_like(x[b:b + 1])
        s64.multiplication(x[b:b + 1].contiguous(), ys)
        assert torch.equal(ys[0], y64[b])
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    o64.rhs_sines(rng.standard_normal((B, 2)), np.array([1, 2], dtype=np.int32), out=f)
    xr = torch.zeros_like(f)
    L.CG.solveHomogen(1e-13, xr, f, o64, stop_rule=L.STOP_REL, maxit=5000)
    its = {}
    for kind in (L.PRECOND_DIAG, L.PRECOND_DEPTH_BLOCK):
        o32.set_precond(kind)
        xm = torch.zeros_like(f)
        info = L.CG.solveMixed(1e-10, xm, f, o64, o32)
        assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-10 for i in info)
        assert float(torch.max(torch.linalg.norm(xm - xr, dim=1) / torch.linalg.norm(xr, dim=1))) < 1e-8
        its[kind] = np.mean([i["iterations"] for i in info])
    print("inner iterations diag %.1f, depth-block %.1f" % (its[0], its[1]), flush=True)
    assert its[L.PRECOND_DEPTH_BLOCK] < 0.7 * its[L.PRECOND_DIAG]


@pytest.mark.parametrize("knob", ["LSG_SP_LDS_BYTES=0", "LSG_SP_SMALL_BYTES=0", "LSG_SP_MEDIUM_BYTES=0",
                                  "LSG_SP_LDS_BYTES=20000", "LSG_SP_IX32=1"])
def test_sparse_scratch_classes_give_the_same_bits(knob):
    """The cell-wise apply launches its items by the size of their point scratch: in LDS with
    four instances per workgroup, in LDS with one (256 or 1024 threads), or in HBM; and it reads
    16-bit copies of its index tables where they fit (LSG_SP_IX32=1: the 32-bit tables).  Every
    variant sums a point's entries in the same order, so operators made with the knob (read when
    the operator is made) must give the default operator's bits, at batch 1 and batch 6, FP64
    and FP32."""
    import torch
    print("scratch classes", knob, flush=True)
    z = np.load(os.path.join(GOLD, "grids", "cfg3_start4_level8.npz"))
    grid = L.AdaptiveSparseGrid(z["idx"], z["active"])
    st = L.VariableCoefficient(L.COEF_CFG3)
    name, val = knob.split("=")
    rng = np.random.default_rng(3)
    for prec, dt in ((64, torch.float64), (32, torch.float32)):
        for B in (1, 6):
            x = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda").to(dt)
            ref = L.MatrixVectorHomogen(grid, st, precision=prec, batch=B)
            yr = torch.zeros_like(x)
            ref.multiplication(x, yr)
            os.environ[name] = val
            try:
                op = L.MatrixVectorHomogen(grid, st, precision=prec, batch=B)
            finally:
                del os.environ[name]
            y = torch.zeros_like(x)
            op.multiplication(x, y)
            assert torch.equal(y, yr), (knob, prec, B)
