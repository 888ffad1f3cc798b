"""The oracle's stored-matrix path (LocalStiffnessMatrices) pinned on the CPU.

Golden vectors come from the compiled reference (tests/golden/make_golden.py, harness
mode `stored`): LocalStiffnessMatricesMemoryDistribution built from the harness's
quadrature plug-in (oracle/ref_plugin.hpp, QuadStencil: a and c vary inside each cell)
and applied by MatrixVectorHomogen::multiplication's StoreLocalStiffnessMatrix branch
(MatrixVectorHomogen.h:170-245,1619-1700), the same plug-in on the fly, and the
preconditioner diagonal psi_p^T A_T psi_p of precondLocalStiffnessMatrix
(cg_method.h:544-641).  The oracle receives the same element matrices, computed here by
quad_element_matrices for its own cells, and must reproduce them to 1e-12.

Beyond the goldens: element matrices of the form a(centre) K + c(centre) M must give the
oracle's on-the-fly operator with those coefficients (and the reference's Ax_var golden),
which pins the entry order independently; a planted mistake in the entry order (p, q
pairs enumerated with q outer) fails the golden; the library's cells are the oracle's.
"""
import json
import os

import numpy as np
import pytest

import loadsg_amd as L
from oracle_ctypes import COEF_GEN, Oracle, gen_coef, quad_element_matrices

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
CASES = ["d3_l4_gen", "d5_l3_gen", "d3_adaptive_gen"]


def load(case):
    base = os.path.join(GOLD, case)
    return {f[:-4]: np.load(os.path.join(base, f)) for f in os.listdir(base) if f.endswith(".npy")}


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def centre_element_matrices(level, pos, coef=gen_coef):
    """a(centre) * Q1 stiffness + c(centre) * Q1 mass (the Poisson / HelmHoltz integrals,
    Stencil.h:873-919, 1707-1752) per cell, in stored-entry order."""
    level = np.asarray(level, dtype=np.int64)
    K, d = level.shape
    h = np.ldexp(1.0, -level)
    a, c = coef((np.asarray(pos) + 0.5) * h)
    E = np.zeros((K, L.cell_entries(d)))
    for e, (p, q) in enumerate(L.cell_pairs(d)):
        same = [((p >> i) & 1) == ((q >> i) & 1) for i in range(d)]
        m1 = [h[:, i] * (1.0 / 3.0 if same[i] else 1.0 / 6.0) for i in range(d)]
        mass = np.prod(m1, axis=0)
        stiff = np.zeros(K)
        for i in range(d):
            k = (1.0 if same[i] else -1.0) / h[:, i]
            stiff += k * np.prod([m1[j] for j in range(d) if j != i], axis=0)
        E[:, e] = a * stiff + c * mass
    return E


@pytest.mark.parametrize("case", CASES)
def test_reference_stored_and_on_the_fly_agree(case):
    gv = load(case)
    assert relmax(gv["Ax_stored"], gv["Ax_stored_otf"]) < 1e-13


@pytest.mark.parametrize("case", CASES)
def test_oracle_stored_matvec_and_diagonal_vs_reference(case):
    gv = load(case)
    o = Oracle(gv["idx"], gv["active"])
    lev, pos = o.cell_boxes()
    o.set_cell_matrices(quad_element_matrices(lev, pos))
    assert relmax(o.matvec(gv["x"]), gv["Ax_stored"]) < 1e-12
    act = gv["active"] == 1
    assert relmax(o.precond()[act], gv["diag_stored"][act]) < 1e-12


@pytest.mark.parametrize("case", CASES)
def test_centre_matrices_reproduce_on_the_fly_operator(case):
    """Stored a(centre) K + c(centre) M equals the coefficient path and the reference's
    Ax_var (the centre-value plug-in) -- independent of the quadrature goldens."""
    gv = load(case)
    o = Oracle(gv["idx"], gv["active"])
    o.set_coef(COEF_GEN)
    y_otf, z_otf = o.matvec(gv["x"]), o.precond()
    lev, pos = o.cell_boxes()
    o.set_cell_matrices(centre_element_matrices(lev, pos))
    y = o.matvec(gv["x"])
    assert relmax(y, y_otf) < 1e-13
    assert relmax(y, gv["Ax_var"]) < 1e-12
    act = gv["active"] == 1
    assert relmax(o.precond()[act], z_otf[act]) < 1e-13


def test_planted_entry_order_mistake_fails_golden():
    """Entries enumerated with q outer, p inner (a plausible transposition of the pair
    order) must not reproduce the reference."""
    gv = load("d3_l4_gen")
    o = Oracle(gv["idx"], gv["active"])
    lev, pos = o.cell_boxes()
    E = quad_element_matrices(lev, pos)
    d = lev.shape[1]
    pairs = L.cell_pairs(d)
    wrong = sorted(range(len(pairs)), key=lambda e: (pairs[e][1], pairs[e][0]))
    o.set_cell_matrices(E[:, wrong])
    assert relmax(o.matvec(gv["x"]), gv["Ax_stored"]) > 1e-3


@pytest.mark.parametrize("case", CASES)
def test_library_cells_are_the_oracles(case):
    """lsg_grid_cell_boxes lists the oracle's cells of depth T+1 (host code, no GPU)."""
    gv = load(case)
    g = L.AdaptiveSparseGrid(gv["idx"], gv["active"])
    lev, pos = g.cell_boxes()
    o = Oracle(gv["idx"], gv["active"])
    olev, opos = o.cell_boxes()
    key = lambda a, b: sorted(map(tuple, np.concatenate([a, b], axis=1).tolist()))
    assert key(lev, pos) == key(olev, opos)
    x, t = g.cells()
    assert np.allclose(x, (pos + 0.5) * np.ldexp(1.0, -lev.astype(np.int64)))
    assert L.cell_entries(lev.shape[1]) == L.lib().lsg_cell_entries(lev.shape[1])
