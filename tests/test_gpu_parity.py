"""GPU parity: the gfx950 path (called through the C ABI) against the compiled
reference's golden vectors and the CPU oracle; size-independent properties at
the full d=5 level-7 shape.

Tolerances: FP64 results within 1e-12 relative (max-norm) of the reference or
oracle (rounding-level differences only: the GPU sums the same terms in a
different order), FP32 within 2e-5; integer data (index sets, iteration counts)
exact.
"""
import json
import os

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import COEF_CFG2, COEF_CFG3, COEF_CONST, COEF_EXPSIN, COEF_GEN, Oracle, gen_coef, regular_grid

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
REGULAR = sorted(c for c in CASES if CASES[c]["argv"][0] == "regular")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


def relmax(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def stencil_of(argv):
    spec = argv[2]
    if spec == "cfg2":
        return L.VariableCoefficient(L.COEF_CFG2), dict(kind=COEF_CFG2)
    if spec == "cfg3":
        return L.VariableCoefficient(L.COEF_CFG3), dict(kind=COEF_CFG3)
    if spec == "gen":     # non-separable a, c: sampled per cell on the host (cell-wise path)
        return (L.CellFunction(a=lambda x: gen_coef(x)[0], c=lambda x: gen_coef(x)[1]), dict(kind=COEF_GEN))
    xi = [float(t) for t in spec[7:].split(",")]
    return L.VariableCoefficient(L.COEF_EXPSIN, xi=xi), dict(kind=COEF_EXPSIN, xi=np.array(xi))


def load(case):
    g = os.path.join(GOLD, case)
    return {f[:-4]: np.load(os.path.join(g, f)) for f in os.listdir(g) if f.endswith(".npy")}


@pytest.mark.parametrize("case", REGULAR)
def test_golden_matvec(case):
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    var, _ = stencil_of(CASES[case]["argv"])
    for st, key in ((L.Poisson(), "Ax_poisson"), (L.HelmHoltz(), "Ax_helm"), (var, "Ax_var")):
        op = L.MatrixVectorHomogen(grid, st)
        y = op.multiplication(gv["x"])
        assert relmax(y, gv[key]) < 1e-12, key


@pytest.mark.parametrize("case", REGULAR)
def test_golden_precond(case):
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    var, _ = stencil_of(CASES[case]["argv"])
    for st, key in ((L.Poisson(), "diag_poisson"), (var, "diag_var")):
        z = L.MatrixVectorHomogen(grid, st).diagonal()
        # the reference evaluates its 1D integrals with cancellation-prone cubic
        # formulas (Stencil.h:757-854), so its diagonal is accurate to ~1e-13 only
        assert relmax(z, gv[key]) < 1e-11, key


@pytest.mark.parametrize("case", REGULAR)
def test_golden_cg(case):
    print("case", case, flush=True)
    gv = load(case)
    grid = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    var, _ = stencil_of(CASES[case]["argv"])
    for st, key in ((L.Poisson(), "cg_poisson"), (var, "cg_var")):
        op = L.MatrixVectorHomogen(grid, st)
        x = np.zeros(grid.size)
        info = L.CG.solveHomogen(1e-10, x, gv["f_rhs"], op, stop_rule=L.STOP_REF_ABS)
        assert info["iterations"] == int(gv[key + "_iters"][0]), key
        assert relmax(x, gv[key + "_x"]) < 1e-10, key


ORACLE_SHAPES = [(2, 8), (3, 6), (4, 5), (5, 4), (6, 3)]


@pytest.mark.parametrize("d,level", ORACLE_SHAPES)
def test_oracle_matvec_and_diag(d, level):
    print("shape", d, level, flush=True)
    idx, act = regular_grid(d, level)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    rng = np.random.default_rng(d * 100 + level)
    xi = rng.normal(0.0, 0.5, 16)
    o = Oracle(idx, act).set_coef(COEF_EXPSIN, xi=xi)
    op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi))
    x = rng.standard_normal(grid.size)
    assert relmax(op.multiplication(x), o.matvec(x)) < 1e-12
    assert relmax(op.diagonal(), o.precond()) < 1e-12
    # FP32 operator: same algorithm in single precision
    op32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi), precision=32)
    assert relmax(op32.multiplication(x), o.matvec(x)) < 2e-5


def test_cfg2_constant_parts_reduce_to_reference_forms():
    """a = 1 + 0.5 prod cos, c = 1 + x_1 against the oracle at d=4 (multi-term path)."""
    d, level = 4, 4
    idx, act = regular_grid(d, level)
    o = Oracle(idx, act).set_coef(COEF_CFG2)
    op = L.MatrixVectorHomogen(L.AdaptiveSparseGrid.regular(d, level), L.VariableCoefficient(L.COEF_CFG2))
    x = np.random.default_rng(3).standard_normal(idx.shape[0])
    assert relmax(op.multiplication(x), o.matvec(x)) < 1e-12
    assert relmax(op.diagonal(), o.precond()) < 1e-12


def test_block_size_invariance_and_repeatability():
    """The autotuner's choice never changes a result: bit-identical outputs."""
    grid = L.AdaptiveSparseGrid.regular(5, 5)
    op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=[0.4, -0.2, 0.3]))
    x = np.random.default_rng(5).standard_normal(grid.size)
    outs = []
    for t in (64, 128, 256, 512):
        op.set_block(t)
        outs.append((t, op.multiplication(x)))
        outs.append((t, op.multiplication(x)))
    for t, y in outs[1:]:
        bad = np.flatnonzero(y != outs[0][1])
        assert bad.size == 0, (t, bad[:10], np.max(np.abs(y - outs[0][1])))


def test_batch_matches_single_instances():
    d, level, B = 4, 4, 3
    grid = L.AdaptiveSparseGrid.regular(d, level)
    rng = np.random.default_rng(11)
    xis = [rng.normal(0, 0.5, 16) for _ in range(B)]
    opb = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), batch=B)
    for b in range(B):
        opb.set_coefficient(b, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b]))
    X = rng.standard_normal((B, grid.size))
    Y = opb.multiplication(X)
    for b in range(B):
        single = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b]))
        assert np.array_equal(Y[b], single.multiplication(X[b]))
    # batched CG == per-instance CG
    F = rng.standard_normal((B, grid.size))
    Xs = np.zeros((B, grid.size))
    infos = L.CG.solveHomogen(1e-10, Xs, F, opb, stop_rule=L.STOP_REL)
    for b in range(B):
        single = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b]))
        xb = np.zeros(grid.size)
        ib = L.CG.solveHomogen(1e-10, xb, F[b], single, stop_rule=L.STOP_REL)
        assert infos[b]["iterations"] == ib["iterations"]
        assert relmax(Xs[b], xb) < 1e-13


def test_full_size_d5_l7_properties():
    """d=5 level 7 (N=61183): symmetry, linearity, positivity, diagonal = e_p^T A e_p."""
    print("d5 l7", flush=True)
    grid = L.AdaptiveSparseGrid.regular(5, 7)
    rng = np.random.default_rng(20261203)
    xi = rng.normal(0.0, 0.5, 16)
    op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xi))
    n = grid.size
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Ax, Ay = op.multiplication(x), op.multiplication(y)
    assert abs(x @ Ay - y @ Ax) <= 1e-12 * abs(x @ Ay)
    Axy = op.multiplication(x + 2.0 * y)
    assert relmax(Axy, Ax + 2.0 * Ay) < 1e-12
    assert x @ Ax > 0 and y @ Ay > 0
    D = op.diagonal()
    for p in rng.choice(n, 6, replace=False):
        e = np.zeros(n)
        e[p] = 1.0
        assert abs(op.multiplication(e)[p] - D[p]) <= 1e-12 * abs(D[p])
    # CG to a relative residual of 1e-8 converges and the residual is what it claims
    f = op.rhs_sines(rng.standard_normal(8), rng.integers(1, 4, 8))
    xs = np.zeros(n)
    info = L.CG.solveHomogen(1e-8, xs, f, op, stop_rule=L.STOP_REL)
    assert info["converged"] == 1
    r = f - op.multiplication(xs)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(f)


def test_full_size_d6_l8_properties():
    """d=6 level 8 (N = 471,041, the largest config-4 shape): FP64 symmetry, linearity,
    diagonal = e_p^T A e_p; FP32 matvec within 2e-5 of FP64; the mixed-precision solve
    reaches a true FP64 relative residual of 1e-8."""
    print("d6 l8", flush=True)
    grid = L.AdaptiveSparseGrid.regular(6, 8)
    n = grid.size
    assert n == 471041
    rng = np.random.default_rng(68)
    st = L.VariableCoefficient(L.COEF_EXPSIN, xi=rng.normal(0.0, 0.5, 16))
    op = L.MatrixVectorHomogen(grid, st)
    op32 = L.MatrixVectorHomogen(grid, st, precision=32)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Ax, Ay = op.multiplication(x), op.multiplication(y)
    assert abs(x @ Ay - y @ Ax) <= 1e-12 * abs(x @ Ay)
    assert relmax(op.multiplication(x - 0.5 * y), Ax - 0.5 * Ay) < 1e-12
    assert x @ Ax > 0
    D = op.diagonal()
    for p in rng.choice(n, 3, replace=False):
        e = np.zeros(n)
        e[p] = 1.0
        assert abs(op.multiplication(e)[p] - D[p]) <= 1e-12 * abs(D[p])
    assert relmax(op32.multiplication(x), Ax) < 2e-5
    f = op.rhs_sines(rng.standard_normal(8), rng.integers(1, 4, 8))
    xs = np.zeros(n)
    info = L.CG.solveMixed(1e-8, xs, f, op, op32, inner_rel_tol=1e-4)
    assert info["converged"] == 1
    r = f - op.multiplication(xs)
    assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(f)


def test_mixed_precision_refinement():
    """FP32 inner PCG under FP64 iterative refinement reaches the FP64 tolerance and
    agrees with the FP64-only solve to the stated error."""
    grid = L.AdaptiveSparseGrid.regular(4, 6)
    rng = np.random.default_rng(8)
    st = L.VariableCoefficient(L.COEF_EXPSIN, xi=rng.normal(0, 0.5, 16))
    op64 = L.MatrixVectorHomogen(grid, st)
    op32 = L.MatrixVectorHomogen(grid, st, precision=32)
    f = op64.rhs_sines(rng.standard_normal(8), rng.integers(1, 4, 8))
    x64 = np.zeros(grid.size)
    L.CG.solveHomogen(1e-12, x64, f, op64, stop_rule=L.STOP_REL)
    xm = np.zeros(grid.size)
    info = L.CG.solveMixed(1e-10, xm, f, op64, op32, inner_rel_tol=1e-4)
    assert info["converged"] == 1 and info["rel_residual"] <= 1e-10
    assert relmax(xm, x64) < 1e-8


def test_rhs_sines_against_quadrature():
    """b_p = int f psi_p against a 6-point Gauss quadrature over the finest cells."""
    d, level = 2, 4
    grid = L.AdaptiveSparseGrid.regular(d, level)
    idx, _ = grid.slot_order()
    op = L.MatrixVectorHomogen(grid, L.Poisson())
    eta, m = np.array([0.7, -1.3]), np.array([1, 3])
    b = op.rhs_sines(eta, m)
    depth = np.floor(np.log2(idx)).astype(int)
    F = int(depth.max())
    gx, gw = np.polynomial.legendre.leggauss(6)
    h = 1.0 / 2 ** F
    xs = ((np.arange(2 ** F)[:, None] + (gx[None, :] + 1) / 2) * h).ravel()
    ws = np.tile(gw * h / 2, 2 ** F)

    def psi1(t, k, x):
        n = 2 ** t
        v = np.zeros(n + 1)
        v[k] = 0.9 if (t >= 2 and k in (1, n - 1)) else 1.0
        if k - 1 >= 1:
            v[k - 1], v[k - 2] = -0.6, 0.1
        if k + 1 <= n - 1:
            v[k + 1], v[k + 2] = -0.6, 0.1
        return np.interp(x, np.arange(n + 1) / n, v)

    ref = np.zeros(grid.size)
    for p in range(grid.size):
        s = 0.0
        for e, mm in zip(eta, m):
            pr = e
            for i in range(d):
                t = depth[p, i]
                k = 2 * (idx[p, i] - 2 ** t) + 1
                pr *= np.sum(ws * np.sin(mm * np.pi * xs) * psi1(t, k, xs))
            s += pr
        ref[p] = s
    assert relmax(b, ref) < 1e-10


def test_family_plan_equals_single_item_plan():
    """Dense grids run on the single-item plan (every prolongation dimension materialised);
    the family plan (last prolongation dimension fused into the apply stage, selected by
    LSG_SINGLE_PLAN_BATCH=0) computes the same operator: FP64 within 1e-13, FP32 within 1e-6,
    at d=5 L5 (dense separable kernel) and d=4 L4 with a multi-term coefficient."""
    import torch
    for d, level, st in ((5, 5, L.VariableCoefficient(L.COEF_EXPSIN, xi=[0.3, -0.2, 0.4, 0.1])),
                         (4, 4, L.VariableCoefficient(L.COEF_CFG2))):
        grid = L.AdaptiveSparseGrid.regular(d, level)
        for prec in (64, 32):
            dt = torch.float64 if prec == 64 else torch.float32
            x = torch.tensor(np.random.default_rng(d).standard_normal((3, grid.size)), dtype=dt, device="cuda")
            ys = {}
            for plan in ("single", "family"):
                old = os.environ.get("LSG_SINGLE_PLAN_BATCH")
                if plan == "family":
                    os.environ["LSG_SINGLE_PLAN_BATCH"] = "0"
                try:
                    op = L.MatrixVectorHomogen(grid, st, precision=prec, batch=3)
                finally:
                    if old is None:
                        os.environ.pop("LSG_SINGLE_PLAN_BATCH", None)
                    else:
                        os.environ["LSG_SINGLE_PLAN_BATCH"] = old
                ys[plan] = torch.zeros_like(x)
                op.multiplication(x, ys[plan])
                ys[plan + "_stats"] = op.stats()
            assert ys["single_stats"]["stream_values"] == ys["family_stats"]["stream_values"]
            err = (torch.max(torch.abs(ys["single"] - ys["family"])) / torch.max(torch.abs(ys["family"]))).item()
            assert err < (1e-13 if prec == 64 else 1e-6), (d, prec, err)
