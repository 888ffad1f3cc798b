"""Discretisation error against a manufactured solution: -Laplace u = f on [0,1]^d with
u = prod_i sin(pi x_i), f = d pi^2 u, zero Dirichlet boundaries.

The energy error of the Galerkin solution u_h follows from orthogonality without
evaluating u_h anywhere: |u - u_h|_a^2 = a(u, u) - a(u_h, u_h) = d pi^2 / 2^d - x . b,
where b_p = int f psi_p is the load vector (lsg_rhs_sines) and x the prewavelet
coefficients of u_h.  On regular sparse grids it decays like 2^-L times a power of L,
so the error roughly halves per level."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("d,levels", [(2, [4, 5, 6, 7, 8]), (3, [3, 4, 5, 6])])
def test_energy_error_decays_with_level(d, levels):
    import torch
    import loadsg_amd as L
    errs = []
    for level in levels:
        print("d=%d level=%d" % (d, level), flush=True)
        grid = L.AdaptiveSparseGrid.regular(d, level)
        op = L.MatrixVectorHomogen(grid, L.Poisson())
        b = torch.zeros(1, grid.size, dtype=torch.float64, device="cuda")
        op.rhs_sines(np.array([[d * math.pi ** 2]]), np.array([1], dtype=np.int32), out=b)
        x = torch.zeros_like(b)
        info = L.CG.solveHomogen(1e-13, x, b, op, stop_rule=L.STOP_REL, maxit=5000)
        assert info["converged"] == 1
        au = d * math.pi ** 2 / 2 ** d
        e2 = au - float((x * b).sum())
        assert e2 > 0.0                       # a(u_h, u_h) < a(u, u) for the Galerkin projection
        errs.append(math.sqrt(e2 / au))
    print("relative energy errors", errs, flush=True)
    for a, b in zip(errs, errs[1:]):
        assert b < a
    assert 1.9 < errs[-2] / errs[-1] < 2.1    # halves per level (measured: 2.000 for d = 2 and 3)
