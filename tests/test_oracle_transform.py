"""Pins the oracle's basis transformations (BasisTransformations.cpp) and the norms
built on them (applications/H1_Norm.cpp) before the device path is checked against them.

1. Golden vectors written by the compiled LoAdSG reference (tests/golden/*/tr_*.npy,
   made by tests/golden/make_golden.py): CalcHierarchicalBasis, calcPrewByNodal,
   calcNodalByPrew, H1::getValue and L2::getValue, and the f_prew vector of the
   reference driver's right-hand side.
2. Brute force that shares no code with the oracle: the prewavelet and hierarchical
   basis functions are evaluated at the sparse-grid points as piecewise-linear
   functions on the finest full grid (E_psi, E_phi).  Nodal values from prewavelets
   must equal E_psi p, prewavelets from nodal values must solve E_psi p = u, and
   hierarchical surpluses must solve E_phi a = u.  A wrong boundary coefficient, a
   dropped depth or a transposed line solve fails these.
3. Known answers of the reference's driver examples_doxygen/poisson_regular.cpp at d=3
   level 7 (iterations and three error norms, as that driver prints them).
"""
import json
import math
import os

import numpy as np
import pytest

from oracle_ctypes import COEF_CONST, Oracle, regular_grid
from test_oracle import interp_to, kron_all, load, prewavelet_nodal_1d, relmax

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLD, "cases.json")))
REGULAR = sorted(c for c in CASES if CASES[c]["argv"][0] == "regular")
ADAPTIVE = sorted(c for c in CASES if CASES[c]["argv"][0].startswith("adaptive"))


def coords(idx):
    depth = np.floor(np.log2(idx)).astype(int)
    k = 2 * (idx - 2 ** depth) + 1
    return k / 2.0 ** depth


@pytest.mark.parametrize("case", REGULAR + ADAPTIVE)
def test_golden_transforms(case):
    """Regular and adaptive grids; on the adaptive ones the prewavelet solve works per
    connected component of each depth's points, with line matrices cut at missing and
    hanging points, and only active points are written."""
    gv = load(case)
    o = Oracle(gv["idx"], gv["active"])
    u = gv["tr_u"]
    assert relmax(o.transform(o.HIER, u), gv["tr_hier"]) < 1e-14
    assert relmax(o.transform(o.PREW_BY_NODAL, u), gv["tr_prew"]) < 1e-13
    assert relmax(o.transform(o.NODAL_BY_PREW, gv["x"]), gv["tr_nodal_of_x"]) < 1e-13
    # the reference driver's f_prew = calcPrewByNodal(d pi^2 prod sin(pi x_i)) at the active points
    d = CASES[case]["d"]
    f = d * math.pi ** 2 * np.prod(np.sin(math.pi * coords(gv["idx"])), axis=1) * gv["active"]
    assert relmax(o.transform(o.PREW_BY_NODAL, f), gv["f_prew"]) < 1e-13


@pytest.mark.parametrize("case", ADAPTIVE)
def test_adaptive_round_trip(case):
    """calcNodalByPrew inverts calcPrewByNodal at the active points, and neither writes a
    hanging point."""
    gv = load(case)
    act = gv["active"].astype(bool)
    assert (~act).any()
    o = Oracle(gv["idx"], gv["active"])
    p = o.transform(o.PREW_BY_NODAL, gv["tr_u"])
    back = o.transform(o.NODAL_BY_PREW, p)
    assert np.abs(back - gv["tr_u"])[act].max() < 1e-13
    assert not p[~act].any() and not back[~act].any()
    assert not o.transform(o.HIER, gv["tr_u"])[~act].any()


@pytest.mark.parametrize("case", REGULAR + ADAPTIVE)
def test_golden_norms(case):
    """H1::getValue / L2::getValue: sqrt(v^T A v) over active points, v = prew(u), with
    the Poisson (H1) or HelmHoltz (L2) operator (H1_Norm.cpp:8-78)."""
    gv = load(case)
    o = Oracle(gv["idx"], gv["active"])
    v = o.transform(o.PREW_BY_NODAL, gv["tr_u"])
    h1 = math.sqrt(v @ o.set_coef(COEF_CONST, a0=1.0, c0=0.0).matvec(v))
    l2 = math.sqrt(v @ o.set_coef(COEF_CONST, a0=0.0, c0=1.0).matvec(v))
    ref = gv["tr_norms"]
    assert abs(h1 - ref[0]) < 1e-12 * ref[0]
    assert abs(l2 - ref[1]) < 1e-12 * ref[1]


def hat_nodal_1d(t, k):
    v = np.zeros(2 ** t + 1)
    v[k] = 1.0
    return v


def basis_at_points(idx, basis):
    """Matrix E[r, q]: basis function of point q evaluated at grid point r."""
    depth = np.floor(np.log2(idx)).astype(int)
    F = int(depth.max())
    n, d = idx.shape
    nF = 2 ** F - 1
    kk = 2 * (idx - 2 ** depth) + 1
    rows = np.zeros(n, dtype=np.int64)
    for i in range(d):
        rows = rows * nF + (kk[:, i] * 2 ** (F - depth[:, i]) - 1)
    cols = []
    for q in range(n):
        vs = [interp_to(basis(depth[q, i], kk[q, i]), depth[q, i], F)[:, None] for i in range(d)]
        cols.append(kron_all(vs)[:, 0][rows])
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("d,L", [(1, 5), (2, 4), (3, 3)])
def test_bruteforce_transforms(d, L):
    idx, act = regular_grid(d, L)
    o = Oracle(idx, act)
    Epsi = basis_at_points(idx, prewavelet_nodal_1d)
    Ephi = basis_at_points(idx, hat_nodal_1d)
    rng = np.random.default_rng(11)
    p = rng.standard_normal(idx.shape[0])
    u = rng.standard_normal(idx.shape[0])
    assert relmax(o.transform(o.NODAL_BY_PREW, p), Epsi @ p) < 1e-13
    assert relmax(o.transform(o.PREW_BY_NODAL, u), np.linalg.solve(Epsi, u)) < 1e-11
    assert relmax(o.transform(o.HIER, u), np.linalg.solve(Ephi, u)) < 1e-12
    assert relmax(o.transform(o.DEHIER, u), Ephi @ u) < 1e-13


def test_round_trip_larger():
    idx, act = regular_grid(3, 6)
    o = Oracle(idx, act)
    u = np.random.default_rng(3).standard_normal(idx.shape[0])
    p = o.transform(o.PREW_BY_NODAL, u)
    assert relmax(o.transform(o.NODAL_BY_PREW, p), u) < 1e-12
    assert relmax(o.transform(o.DEHIER, o.transform(o.HIER, u)), u) < 1e-13


# Known answers printed by the reference's own driver examples_doxygen/poisson_regular.cpp
# (d=3, level 7): CG iterations, L_infty(u - u_h), L2::getValue(u - u_h) and
# L2 + H1::getValue(u - u_h), six significant digits.
POISSON_D3_L7 = dict(iterations=21, linf=2.98384e-4, l2=2.34495e-5, h1_plus_l2=5.17872e-3)


def test_poisson_driver_known_answers():
    """The whole driver pipeline (poisson_regular.cpp:57-147) on the oracle: f_h =
    M prew(f_I), CG to eps 1e-10 with the reference stop rule, u_h = calcNodalByPrew(x),
    errors of u - u_h in the reference's norms."""
    d, L = 3, 7
    idx, act = regular_grid(d, L)
    u = np.prod(np.sin(math.pi * coords(idx)), axis=1)
    o = Oracle(idx, act)
    p = o.transform(o.PREW_BY_NODAL, d * math.pi ** 2 * u)
    f = o.set_coef(COEF_CONST, a0=0.0, c0=1.0).matvec(p)
    x, it, _ = o.set_coef(COEF_CONST, a0=1.0, c0=0.0).cg(f, eps=1e-10, maxit=500)
    e = u - o.transform(o.NODAL_BY_PREW, x)
    pe = o.transform(o.PREW_BY_NODAL, e)
    l2 = math.sqrt(pe @ o.set_coef(COEF_CONST, a0=0.0, c0=1.0).matvec(pe))
    h1 = math.sqrt(pe @ o.set_coef(COEF_CONST, a0=1.0, c0=0.0).matvec(pe))
    k = POISSON_D3_L7
    assert it == k["iterations"]
    assert abs(np.abs(e).max() - k["linf"]) < 1e-5 * k["linf"]
    assert abs(l2 - k["l2"]) < 1e-5 * k["l2"]
    assert abs(l2 + h1 - k["h1_plus_l2"]) < 1e-5 * k["h1_plus_l2"]
