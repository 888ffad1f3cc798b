"""Instances on grids of different dimension and level solved together
(loadsg_amd.solve_many): every instance converges and matches its own solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _group(L, d, level, xis, etas, m):
    import torch
    B = len(xis)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    op64 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=32, batch=B)
    for b in range(B):
        c = L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b])
        op64.set_coefficient(b, c)
        op32.set_coefficient(b, c)
    f = torch.zeros(B, grid.size, dtype=torch.float64, device="cuda")
    op64.rhs_sines(np.asarray(etas), m, out=f)
    return torch.zeros_like(f), f, op64, op32


def test_solve_many_matches_single_solves():
    import loadsg_amd as L
    rng = np.random.default_rng(8)
    m = np.array([1, 2, 3], dtype=np.int32)
    shapes = [(2, 5, 3), (3, 4, 2), (4, 3, 1), (3, 5, 2)]     # (d, level, instances)
    probs, single = [], []
    for d, level, B in shapes:
        xis = rng.normal(0.0, 0.5, (B, 6))
        etas = rng.standard_normal((B, 3))
        print("group d=%d level=%d B=%d" % (d, level, B), flush=True)
        probs.append(_group(L, d, level, xis, etas, m))
        single.append([_group(L, d, level, xis[b:b + 1], etas[b:b + 1], m) for b in range(B)])
    infos = L.solve_many(probs, 1e-8)
    for (x, f, _, _), inf, sgl in zip(probs, infos, single):
        assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-8 for i in inf)
        for b, (xs, fs, o64, o32) in enumerate(sgl):
            i1 = L.CG.solveMixed(1e-8, xs, fs, o64, o32)
            assert i1["iterations"] == inf[b]["iterations"]
            assert float((xs[0] - x[b]).abs().max()) <= 1e-12 * float(x[b].abs().max())


def test_uneven_batch_config4_shapes():
    """16 instances with d in {3..6}, level in {5..7} (the uneven-batch workload at
    reduced count), exp-sine coefficients and 8-term sine right-hand sides: every
    instance reaches ||f - A x|| <= 1e-8 ||f|| (checked here in FP64)."""
    import torch
    import loadsg_amd as L
    rng = np.random.default_rng(20261204)
    ds = rng.integers(3, 7, 16)
    ls = rng.integers(5, 8, 16)
    m = rng.integers(1, 4, 8).astype(np.int32)
    groups = {}
    for d, level in zip(ds, ls):
        groups.setdefault((int(d), int(level)), 0)
        groups[(int(d), int(level))] += 1
    probs = []
    for (d, level), B in sorted(groups.items()):
        print("group d=%d level=%d B=%d" % (d, level, B), flush=True)
        probs.append(_group(L, d, level, rng.normal(0.0, 0.5, (B, 16)), rng.standard_normal((B, 8)), m))
    infos = L.solve_many(probs, 1e-8)
    for (x, f, op64, _), inf in zip(probs, infos):
        assert all(i["converged"] == 1 for i in inf)
        Ax = torch.zeros_like(f)
        op64.multiplication(x, Ax)
        rel = torch.linalg.norm(f - Ax, dim=1) / torch.linalg.norm(f, dim=1)
        assert float(rel.max()) <= 1e-8


def test_full_config4_matches_single_solves():
    """The uneven batch at its full size: 64 instances, d uniform in {3..6}, level in {5..8}
    (the bench's seeded draw, 15 grid shapes up to d=6 level 8, N = 471,041), exp-sine
    coefficients with 16 xi and 8-term sine right-hand sides, solved together by
    lsg_ir_solve_many to 1e-8; every instance equals its own single-instance solve
    (inner iterations and refinement sweeps equal, x within 1e-12)."""
    import torch
    import loadsg_amd as L
    SEED = 20261203
    rng = np.random.default_rng(SEED + 4)
    B = 64
    ds = rng.integers(3, 7, B)
    ls = rng.integers(5, 9, B)
    xi = rng.normal(0.0, 0.5, (B, 16))
    eta = rng.standard_normal((B, 8))
    m = rng.integers(1, 4, 8).astype(np.int32)
    groups = {}
    for b in range(B):
        groups.setdefault((int(ds[b]), int(ls[b])), []).append(b)
    assert max(ls) == 8 and len(groups) >= 12
    probs, members = [], []
    for (d, level), idx in sorted(groups.items()):
        print("group d=%d level=%d B=%d" % (d, level, len(idx)), flush=True)
        probs.append(_group(L, d, level, xi[idx], eta[idx], m))
        members.append((d, level, idx))
    infos = L.solve_many(probs, 1e-8)
    for (x, f, op64, _), inf, (d, level, idx) in zip(probs, infos, members):
        assert all(i["converged"] == 1 and i["rel_residual"] <= 1e-8 for i in inf)
        for k, b in enumerate(idx):
            xs, fs, o64, o32 = _group(L, d, level, xi[b:b + 1], eta[b:b + 1], m)
            i1 = L.CG.solveMixed(1e-8, xs, fs, o64, o32)
            assert i1["iterations"] == inf[k]["iterations"] and i1["outer"] == inf[k]["outer"], (d, level, b)
            assert float((xs[0] - x[k]).abs().max()) <= 1e-12 * float(x[k].abs().max())
