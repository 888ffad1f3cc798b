"""The FP32 prolongation stage loads whole chains into registers; its launch shape
depends on the longest chain of the grid (128, 256 or 512 threads, 16 to 24 values per
thread at these sizes).  Each shape is run here at a grid size that selects it:
d=2 L10 (512 threads), d=3 L8 (128 threads, 22 values per thread) against the oracle,
d=4 L8 (256 threads) and d=5 L7 (128 threads, 17 per thread) against the FP64 operator
(whose prolongation kernel is the per-box one).  Tolerances: FP64 1e-12, FP32 2e-5
(max-norm relative, as in test_gpu_parity) against the oracle, 5e-5 against FP64 at
the two large sizes (batch 2, each instance with its own coefficient)."""
import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import COEF_EXPSIN, Oracle, regular_grid
from test_gpu_parity import relmax

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L.lib()
    yield


@pytest.mark.parametrize("d,level", [(2, 10), (3, 8)])
def test_fp32_chain_shapes_vs_oracle(d, level):
    print("shape", d, level, flush=True)
    idx, act = regular_grid(d, level)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    rng = np.random.default_rng(40 + d)
    xi = rng.normal(0.0, 0.5, 16)
    x = rng.standard_normal(grid.size)
    y_o = Oracle(idx, act).set_coef(COEF_EXPSIN, xi=xi).matvec(x)
    st = L.VariableCoefficient(L.COEF_EXPSIN, xi=xi)
    assert relmax(L.MatrixVectorHomogen(grid, st).multiplication(x), y_o) < 1e-12
    assert relmax(L.MatrixVectorHomogen(grid, st, precision=32).multiplication(x), y_o) < 2e-5


@pytest.mark.parametrize("d,level", [(4, 8), (5, 7)])
def test_fp32_chain_shapes_vs_fp64(d, level):
    import torch
    print("shape", d, level, flush=True)
    grid = L.AdaptiveSparseGrid.regular(d, level)
    rng = np.random.default_rng(50 + d)
    B = 2
    xis = rng.normal(0.0, 0.5, (B, 16))
    op64 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=64, batch=B)
    op32 = L.MatrixVectorHomogen(grid, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[0]), precision=32, batch=B)
    op64.set_coefficient(1, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[1]))
    op32.set_coefficient(1, L.VariableCoefficient(L.COEF_EXPSIN, xi=xis[1]))
    x = torch.tensor(rng.standard_normal((B, grid.size)), device="cuda")
    y64 = torch.zeros_like(x)
    op64.multiplication(x, y64)
    x32 = x.float()
    y32 = torch.zeros_like(x32)
    op32.multiplication(x32, y32)
    for b in range(B):
        assert relmax(y32[b].double().cpu().numpy(), y64[b].cpu().numpy()) < 5e-5, b
