"""Manufactured solutions of the benchmark problems and their discretisation errors.

-div(a grad u) + c u = f on (0,1)^d with zero Dirichlet boundaries, u given in closed
form.  a, c and u are written as sums of products of 1D functions; f and the integrals
the error needs then are sums of products too:

  f     = sum_{m,i} -(alpha_mi' v_i' + alpha_mi v_i'') prod_{j!=i} alpha_mj v_j
          + sum_m prod_j gamma_mj v_j                  (a = sum_m prod alpha_mj, c = sum_m prod gamma_mj)
  |u|_a^2 = int a |grad u|^2 + c u^2,   ||u||^2 = prod_j int v_j^2.

The load vector b_p = int f psi_p comes from the library's separable load vector
(exact 1D moments on the device side), so x solves the Galerkin system and

  L2 error      ||u - u_h||^2 = ||u||^2 - 2 x . b_u + x^T M x,   b_u = (u, psi_p),
                M the mass (HelmHoltz) operator in the prewavelet basis (the exact Gram
                matrix of the prewavelets on a regular grid)
  energy error  |u - u_h|_a^2 = |u|_a^2 - x . b      (Galerkin orthogonality; constant a
                and c only: a variable coefficient enters the element matrices through its
                cell-centre value)

hold without evaluating u_h off the grid.  The nodal error max_p |u(x_p) - u_h(x_p)| over
the grid points uses calcNodalByPrew (BasisTransformations.cpp:670-707), the quantity the
reference's driver prints as L_infty (examples_doxygen/poisson_regular.cpp:130-150).
The reference's own right-hand side convention f_h = M prew(f_I) (poisson_regular.cpp:
57-113) is available as reference_rhs().

Solutions that are not sums of products (config 3: u = |x|^(2/3) prod x_i (1 - x_i),
a = 1 + |x|^2 on a grid refined towards the origin, where grad u is singular) are given
pointwise (PointwiseProblem).  Their right-hand side is the reference driver's
f_h = M prew(f_I), and their errors are the ones its drivers print: the largest nodal
error over the active points and the L2 / H1 norms (L2::getValue, H1::getValue,
applications/H1_Norm.cpp:8-78) of the sparse-grid interpolant of the nodal error.
"""
import numpy as np

from . import (CG, CG_NO_PRECOND, COEF_CFG2, COEF_CFG3, H1, L2, STOP_REL, HelmHoltz, MatrixVectorHomogen,
               Poisson, VariableCoefficient)

_GX, _GW = np.polynomial.legendre.leggauss(40)


def integrate1d(g, pieces=16):
    """int_0^1 g(x) dx by a 40-point Gauss rule on `pieces` equal intervals (g smooth)."""
    tot = 0.0
    for k in range(pieces):
        x = (k + 0.5 * (1.0 + _GX)) / pieces
        tot += np.dot(_GW, g(x)) * 0.5 / pieces
    return float(tot)


def coordinates(grid):
    """Coordinates of the grid points in slot order: 1D index v of depth t = floor(log2 v)
    sits at (2 (v - 2^t) + 1) / 2^t (indices/index.h:718-730)."""
    idx, _ = grid.slot_order()
    t = np.floor(np.log2(idx)).astype(np.int64)
    return (2.0 * (idx - 2.0 ** t) + 1.0) / 2.0 ** t


class Fn1:
    """1D function with its first two derivatives (vectorised callables)."""

    def __init__(self, f, df, d2f=None):
        self.f, self.df, self.d2f = f, df, d2f


def const1(v):
    return Fn1(lambda x: np.full_like(np.asarray(x, dtype=float), v), lambda x: np.zeros_like(np.asarray(x, float)),
               lambda x: np.zeros_like(np.asarray(x, float)))


def _mul(*fs):
    return lambda x: np.prod([f(x) for f in fs], axis=0)


class Problem:
    """u = prod_j v_j, a = sum_m prod_j alpha[m][j], c = sum_m prod_j gamma[m][j]."""

    def __init__(self, name, d, v, alpha, gamma, stencil):
        self.name, self.d, self.v, self.alpha, self.gamma, self.stencil = name, d, v, alpha, gamma, stencil

    # ---- pointwise values
    def u(self, x):
        return np.prod([self.v[j].f(x[:, j]) for j in range(self.d)], axis=0)

    def f(self, x):
        return sum(np.prod([fac(x[:, j]) for j, fac in enumerate(t)], axis=0) for t in self.f_terms())

    # ---- separable decompositions
    def f_terms(self):
        d, v = self.d, self.v
        terms = []
        for al in self.alpha:
            for i in range(d):
                fi1 = (lambda a, w: lambda x: -a.df(x) * w.df(x))(al[i], v[i])
                fi2 = (lambda a, w: lambda x: -a.f(x) * w.d2f(x))(al[i], v[i])
                rest = [(lambda a, w: _mul(a.f, w.f))(al[j], v[j]) for j in range(d)]
                terms.append([fi1 if j == i else rest[j] for j in range(d)])
                terms.append([fi2 if j == i else rest[j] for j in range(d)])
        for ga in self.gamma:
            terms.append([(lambda g, w: _mul(g.f, w.f))(ga[j], v[j]) for j in range(d)])
        return terms

    def u_terms(self):
        return [[self.v[j].f for j in range(self.d)]]

    def energy_sq(self):
        """|u|_a^2 = int a |grad u|^2 + c u^2 as sums of products of 1D integrals."""
        d, v = self.d, self.v
        tot = 0.0
        for al in self.alpha:
            for i in range(d):
                p = 1.0
                for j in range(d):
                    w = v[j].df if j == i else v[j].f
                    p *= integrate1d((lambda a, w: lambda x: a.f(x) * w(x) ** 2)(al[j], w))
                tot += p
        for ga in self.gamma:
            p = 1.0
            for j in range(d):
                p *= integrate1d((lambda g, w: lambda x: g.f(x) * w.f(x) ** 2)(ga[j], v[j]))
            tot += p
        return tot

    def l2_sq(self):
        return float(np.prod([integrate1d((lambda w: lambda x: w.f(x) ** 2)(self.v[j])) for j in range(self.d)]))

    # ---- discrete quantities
    def load(self, op):
        """b_p = int f psi_p (host slot order)."""
        return op.rhs_separable(self.f_terms())

    def reference_rhs(self, grid, op_fp64):
        """The reference driver's f_h = M prew(f_I): nodal values of f at the active points,
        calcPrewByNodal, then the mass operator (poisson_regular.cpp:57-113)."""
        _, act = grid.slot_order()
        fI = self.f(coordinates(grid)) * act
        mass = MatrixVectorHomogen(grid, HelmHoltz())
        return mass.multiplication(op_fp64.prew_by_nodal(fI))

    def errors(self, grid, x, b, op):
        """Discretisation errors of the prewavelet coefficients x (slot order) of the Galerkin
        solution of A x = b with b = load(op)."""
        mass = MatrixVectorHomogen(grid, HelmHoltz())
        bu = op.rhs_separable(self.u_terms())
        l2 = self.l2_sq() - 2.0 * float(np.dot(x, bu)) + float(np.dot(x, mass.multiplication(x)))
        _, act = grid.slot_order()
        uh = op.nodal_by_prew(x)
        nod = np.abs(self.u(coordinates(grid)) - uh)[act == 1]
        out = {"l2": float(np.sqrt(max(l2, 0.0))), "nodal_max": float(nod.max()), "l2_sq_raw": l2}
        if self.exact_galerkin():
            e2 = self.energy_sq() - float(np.dot(x, b))
            out.update(energy=float(np.sqrt(max(e2, 0.0))), energy_sq_raw=e2)
        return out

    def exact_galerkin(self):
        """True when the operator is the exact Galerkin matrix (constant a and c): the element
        matrices of a variable coefficient use its cell-centre value (a midpoint rule), so
        the energy identity holds only up to that quadrature error."""
        def constant(terms):
            return all(len(t) == 0 or all(np.allclose(g.df(np.linspace(0.1, 0.9, 5)), 0.0) for g in t)
                       for t in terms)
        return constant(self.alpha) and constant(self.gamma)

    def solve(self, grid, tol=1e-10, precondition=True, maxit=5000):
        """FP64 CG to ||f - A x|| <= tol ||f||; returns (x, b, info, op)."""
        op = MatrixVectorHomogen(grid, self.stencil)
        b = self.load(op)
        x = np.zeros(grid.size)
        rule = STOP_REL if precondition else (STOP_REL | CG_NO_PRECOND)
        info = CG.solveHomogen(tol, x, b, op, stop_rule=rule, maxit=maxit)
        return x, b, info, op


def cfg1(d=2):
    """-Laplace u = f, u = prod sin(pi x_i), f = d pi^2 u (Poisson)."""
    s = Fn1(lambda x: np.sin(np.pi * x), lambda x: np.pi * np.cos(np.pi * x), lambda x: -np.pi ** 2 * np.sin(np.pi * x))
    return Problem("cfg1", d, [s] * d, [[const1(1.0)] * d], [], Poisson())


def cfg2(d=4):
    """-div(a grad u) + c u = f, a = 1 + 0.5 prod cos(pi x_i), c = 1 + x_1,
    u = prod x_i (1 - x_i) * exp(x_1)."""
    v1 = Fn1(lambda x: x * (1 - x) * np.exp(x), lambda x: np.exp(x) * (1 - x - x * x),
             lambda x: -np.exp(x) * x * (3 + x))
    vj = Fn1(lambda x: x * (1 - x), lambda x: 1 - 2 * x, lambda x: np.full_like(np.asarray(x, float), -2.0))
    cos1 = Fn1(lambda x: 0.5 * np.cos(np.pi * x), lambda x: -0.5 * np.pi * np.sin(np.pi * x))
    cosj = Fn1(lambda x: np.cos(np.pi * x), lambda x: -np.pi * np.sin(np.pi * x))
    xfn = Fn1(lambda x: np.asarray(x, float), lambda x: np.ones_like(np.asarray(x, float)))
    alpha = [[const1(1.0)] * d, [cos1] + [cosj] * (d - 1)]
    gamma = [[const1(1.0)] * d, [xfn] + [const1(1.0)] * (d - 1)]
    return Problem("cfg2", d, [v1] + [vj] * (d - 1), alpha, gamma, VariableCoefficient(COEF_CFG2))


class PointwiseProblem:
    """-div(a grad u) + c u = f with u and f given pointwise on points x (n x d)."""

    def __init__(self, name, d, u, f, stencil):
        self.name, self.d, self.u, self.f, self.stencil = name, d, u, f, stencil

    def reference_rhs(self, grid, op_fp64):
        """f_h = M prew(f_I): f at the active points, calcPrewByNodal, the mass operator
        (host slot order)."""
        _, act = grid.slot_order()
        fI = self.f(coordinates(grid)) * act
        mass = MatrixVectorHomogen(grid, HelmHoltz())
        return mass.multiplication(op_fp64.prew_by_nodal(fI))

    def nodal_error(self, grid, x, op_fp64):
        """u(x_p) - u_h(x_p) at the active points, 0 at the hanging ones (slot order)."""
        _, act = grid.slot_order()
        return (self.u(coordinates(grid)) - op_fp64.nodal_by_prew(x)) * act

    def errors(self, grid, x, op_fp64):
        e = self.nodal_error(grid, x, op_fp64)
        return {"nodal_max": float(np.abs(e).max()), "l2": float(L2(grid).getValue(e)),
                "h1": float(H1(grid).getValue(e))}

    def solve(self, grid, tol=1e-10, maxit=5000):
        """FP64 PCG on the reference right-hand side; returns (x, f, info, op)."""
        op = MatrixVectorHomogen(grid, self.stencil)
        f = self.reference_rhs(grid, op)
        x = np.zeros(grid.size)
        info = CG.solveHomogen(tol, x, f, op, stop_rule=STOP_REL, maxit=maxit)
        return x, f, info, op


def _cfg3_parts(x):
    """r, g = r^(2/3), grad g, lap g, P = prod x_i (1 - x_i), grad P, lap P."""
    x = np.atleast_2d(np.asarray(x, dtype=float))
    d = x.shape[1]
    r = np.sqrt(np.sum(x * x, axis=1))
    g = r ** (2.0 / 3.0)
    grad_g = (2.0 / 3.0) * r[:, None] ** (-4.0 / 3.0) * x
    lap_g = (2.0 / 3.0) * (2.0 / 3.0 + d - 2.0) * r ** (-4.0 / 3.0)
    q = x * (1.0 - x)
    P = np.prod(q, axis=1)
    others = np.stack([np.prod(np.delete(q, i, axis=1), axis=1) for i in range(d)], axis=1)
    grad_P = (1.0 - 2.0 * x) * others
    lap_P = -2.0 * np.sum(others, axis=1)
    return x, r, g, grad_g, lap_g, P, grad_P, lap_P


def cfg3(d=3):
    """-div(a grad u) = f on (0,1)^d, a = 1 + |x|^2, u = |x|^(2/3) prod x_i (1 - x_i):
    f = -(grad a . grad u + a lap u) with grad a = 2x, lap |x|^a = a (a + d - 2) |x|^(a-2)."""
    def u(x):
        _, _, g, _, _, P, _, _ = _cfg3_parts(x)
        return g * P

    def f(x):
        x, r, g, grad_g, lap_g, P, grad_P, lap_P = _cfg3_parts(x)
        grad_u = P[:, None] * grad_g + g[:, None] * grad_P
        lap_u = P * lap_g + 2.0 * np.sum(grad_g * grad_P, axis=1) + g * lap_P
        return -(np.sum(2.0 * x * grad_u, axis=1) + (1.0 + r * r) * lap_u)

    return PointwiseProblem("cfg3", d, u, f, VariableCoefficient(COEF_CFG3))
