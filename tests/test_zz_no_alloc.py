"""lsg_cg_solve_noalloc: the CG solve whose work buffers (CG vectors, diagonal, partial
sums, per-instance state) live in the library's fixed device scratch instead of memory
it allocates.  Checked through the C ABI against lsg_cg_solve on the same operator and
inputs: iteration counts equal and solutions within 1e-12 (relative, max-norm), on two
calls in a row, the second on a smaller problem; sizes beyond the scratch are rejected
with LSG_ERR_CAPACITY."""
import ctypes

import numpy as np
import pytest

import loadsg_amd as L

pytestmark = pytest.mark.gpu

LSG_ERR_CAPACITY = -3


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = L.lib()
    f = lib.lsg_cg_solve_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p]
    yield


def _problem(d, level, batch, precision, seed):
    import torch
    rng = np.random.default_rng(seed)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    xis = rng.normal(0.0, 0.5, (batch, 8))
    op = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=precision,
                               batch=batch)
    for b in range(1, batch):
        op.set_coefficient(b, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[b]))
    dt = torch.float64 if precision == 64 else torch.float32
    f = torch.zeros(batch, grid.size, dtype=dt, device="cuda")
    op.rhs_sines(rng.standard_normal((batch, 3)), np.array([1, 2, 3], dtype=np.int32), out=f)
    return grid, op, f


def _solve(fn, op, f, eps, rule, maxit):
    import torch
    x = torch.zeros_like(f)
    info = (L.SolveInfo * op.batch)()
    torch.cuda.synchronize()
    if fn == "noalloc":
        rc = L.lib().lsg_cg_solve_noalloc(op._h, L._ptr(f), L._ptr(x), eps, rule, maxit, info)
    else:
        rc = L.lib().lsg_cg_solve(op._h, L._ptr(f), L._ptr(x), eps, rule, maxit, info, L.LSG_DEVICE)
    return rc, x, [info[b].as_dict() for b in range(op.batch)]


@pytest.mark.parametrize("precision", [64, 32])
def test_noalloc_matches_cg_solve_twice(precision):
    import torch
    cases = [(3, 5, 3, L.STOP_REF_ABS, 1e-10), (2, 4, 2, L.STOP_REL | L.CG_NO_PRECOND, 1e-6)]
    for k, (d, level, batch, rule, eps) in enumerate(cases):       # the second call is the smaller one
        print("noalloc case", precision, d, level, batch, rule, flush=True)
        grid, op, f = _problem(d, level, batch, precision, seed=10 + k)
        rc0, x0, i0 = _solve("cg", op, f, eps, rule, 2000)
        rc1, x1, i1 = _solve("noalloc", op, f, eps, rule, 2000)
        assert rc0 == 0 and rc1 == 0
        assert [i["iterations"] for i in i1] == [i["iterations"] for i in i0]
        assert all(i["converged"] == 1 for i in i1)
        diff = torch.max(torch.abs(x1 - x0)).item() / torch.max(torch.abs(x0)).item()
        assert diff < (1e-12 if precision == 64 else 1e-6), diff


def test_noalloc_rejects_sizes_beyond_its_scratch():
    grid = L.AdaptiveSparseGrid.regular(5, 7)
    op = L.MatrixVectorHomogen(grid, L.Poisson(), precision=64, batch=28)   # 28 x 61,183 x 5 doubles > 64 MiB
    import torch
    f = torch.ones(28, grid.size, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(f)
    info = (L.SolveInfo * 28)()
    rc = L.lib().lsg_cg_solve_noalloc(op._h, L._ptr(f), L._ptr(x), 1e-8, L.STOP_REL, 10, info)
    assert rc == LSG_ERR_CAPACITY
    assert b"scratch" in L.lib().lsg_last_error()
