"""CPU checks of the manufactured-solution module (loadsg_amd/manufactured.py): the
closed-form right-hand sides f = -div(a grad u) + c u against second-order finite
differences of the PDE, and the separable integrals |u|_a^2 and ||u||^2 against a
direct tensor Gauss quadrature of the pointwise integrands (gradients by central
differences), so a dropped term or a wrong sign in the separable decomposition fails."""
import itertools

import numpy as np
import pytest

from loadsg_amd import manufactured as M

CASES = [
    ("cfg1", lambda: M.cfg1(2), lambda x: np.ones(len(x)), lambda x: np.zeros(len(x))),
    ("cfg1_d3", lambda: M.cfg1(3), lambda x: np.ones(len(x)), lambda x: np.zeros(len(x))),
    ("cfg2", lambda: M.cfg2(4), lambda x: 1 + 0.5 * np.prod(np.cos(np.pi * x), axis=1), lambda x: 1 + x[:, 0]),
    ("cfg2_d2", lambda: M.cfg2(2), lambda x: 1 + 0.5 * np.prod(np.cos(np.pi * x), axis=1), lambda x: 1 + x[:, 0]),
]


def fd_rhs(p, a, c, x, h):
    d = x.shape[1]
    f = c(x) * p.u(x)
    for i in range(d):
        e = np.zeros(d)
        e[i] = h / 2

        def flux(y):
            return a(y) * (p.u(y + e) - p.u(y - e)) / h
        f -= (flux(x + e) - flux(x - e)) / h
    return f


@pytest.mark.parametrize("name,mk,a,c", CASES)
def test_rhs_matches_finite_differences(name, mk, a, c):
    p = mk()
    x = np.random.default_rng(1).uniform(0.05, 0.95, (20, p.d))
    e1 = np.max(np.abs(fd_rhs(p, a, c, x, 1e-2) - p.f(x)))
    e2 = np.max(np.abs(fd_rhs(p, a, c, x, 1e-3) - p.f(x)))
    assert e2 < 2e-4 * np.max(np.abs(p.f(x)))
    assert e1 / e2 > 50          # second order: error drops by ~100 per decade of h


@pytest.mark.parametrize("name,mk,a,c", CASES)
def test_separable_integrals_match_tensor_quadrature(name, mk, a, c):
    p = mk()
    d = p.d
    gx, gw = np.polynomial.legendre.leggauss(14 if d <= 2 else 8)
    # composite rule: 2 pieces per dimension
    xs = np.concatenate([(k + 0.5 * (1 + gx)) / 2 for k in range(2)])
    ws = np.concatenate([gw / 4 for _ in range(2)])
    pts = np.array(list(itertools.product(xs, repeat=d)))
    wts = np.prod(np.array(list(itertools.product(ws, repeat=d))), axis=1)
    h = 1e-5
    grad2 = np.zeros(len(pts))
    for i in range(d):
        e = np.zeros(d)
        e[i] = h
        grad2 += ((p.u(pts + e) - p.u(pts - e)) / (2 * h)) ** 2
    u = p.u(pts)
    energy = np.dot(wts, a(pts) * grad2 + c(pts) * u * u)
    l2 = np.dot(wts, u * u)
    assert abs(p.energy_sq() - energy) < 1e-7 * energy
    assert abs(p.l2_sq() - l2) < 1e-10 * l2


@pytest.mark.parametrize("d", [2, 3])
def test_cfg3_rhs_matches_finite_differences(d):
    """Config 3 (a = 1 + |x|^2, u = |x|^(2/3) prod x_i (1 - x_i)): f against central
    differences of the flux a grad u at random interior points (O(h^2) = 1e-8); a dropped
    grad a . grad u term, a wrong Laplacian of |x|^(2/3) (its d-dependence) or a sign
    fails this."""
    P = M.cfg3(d)
    x = np.random.default_rng(d).uniform(0.05, 0.95, (30, d))
    h = 1e-4

    def a(y):
        return 1.0 + np.sum(y * y, axis=1)

    div = np.zeros(len(x))
    for i in range(d):
        e = np.zeros(d)
        e[i] = h
        div += (a(x + e / 2) * (P.u(x + e) - P.u(x)) - a(x - e / 2) * (P.u(x) - P.u(x - e))) / h ** 2
    f = P.f(x)
    assert np.max(np.abs(-div - f) / np.abs(f)) < 1e-6
