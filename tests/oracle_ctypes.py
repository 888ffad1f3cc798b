"""ctypes binding of the CPU oracle (oracle/lsg_oracle.cpp).

Test infrastructure only: the oracle is the checker, never the thing measured
or shipped.  Build it with ``make -C oracle oracle``.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

COEF_CONST, COEF_CFG2, COEF_CFG3, COEF_EXPSIN = 0, 1, 2, 3
# non-separable pair a = 1 + x1 x2 + sin(pi (x1 + x2)), c = 1 + sin(pi x1 x2) / 2 (x2 = x1 at d = 1)
COEF_GEN = 10


def gen_coef(x):
    """(a, c) of COEF_GEN at the points x (rows of d coordinates)."""
    x = np.atleast_2d(x)
    x1, x2 = x[:, 0], x[:, 1 % x.shape[1]]
    return 1.0 + x1 * x2 + np.sin(np.pi * (x1 + x2)), 1.0 + 0.5 * np.sin(np.pi * x1 * x2)


def quad_element_matrices(level, pos, coef=gen_coef):
    """Element matrices of the harness's quadrature plug-in (oracle/ref_plugin.hpp,
    QuadStencil): entry (p, q) of cell k = int_cell a grad phi_p . grad phi_q + c phi_p phi_q
    by the tensor 2-point Gauss-Legendre rule, for the cells prod_i [pos_i 2^-level_i,
    (pos_i + 1) 2^-level_i] (arrays ncells x d).  Returns ncells x 2^(d-1) (2^d + 1) entries
    in the order of the nested CellIndexIterator loops (corners p <= q, bit i = upper end)."""
    level = np.asarray(level, dtype=np.int64)
    pos = np.asarray(pos, dtype=np.float64)
    K, d = level.shape
    h = np.ldexp(1.0, -level)
    lo = pos * h
    g = 0.5 / np.sqrt(3.0)
    nc = 1 << d
    pairs = [(p, q) for p in range(nc) for q in range(p, nc)]
    E = np.zeros((K, len(pairs)))
    for m in range(nc):
        t = np.array([0.5 + g if (m >> i) & 1 else 0.5 - g for i in range(d)])
        x = lo + t * h
        a, c = coef(x)
        w = np.ones(K)
        for i in range(d):
            w = w * (0.5 * h[:, i])
        for e, (p, q) in enumerate(pairs):
            fp = [t[i] if (p >> i) & 1 else 1.0 - t[i] for i in range(d)]
            fq = [t[i] if (q >> i) & 1 else 1.0 - t[i] for i in range(d)]
            mass = 1.0
            for i in range(d):
                mass = mass * (fp[i] * fq[i])
            stiff = np.zeros(K)
            for i in range(d):
                prod = (1.0 if ((p >> i) & 1) == ((q >> i) & 1) else -1.0) / (h[:, i] * h[:, i])
                for j in range(d):
                    if j != i:
                        prod = prod * (fp[j] * fq[j])
                stiff = stiff + prod
            E[:, e] += w * (a * stiff + c * mass)
    return E


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_regular_grid.restype = ctypes.c_int64
        L.orc_regular_grid.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_coef.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_void_p, ctypes.c_int]
        L.orc_num_depths.argtypes = [ctypes.c_void_p]
        L.orc_depths.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_multilevel_size.restype = ctypes.c_int64
        L.orc_multilevel_size.argtypes = [ctypes.c_void_p]
        L.orc_sorted_depths.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.orc_matvec.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_precond.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_cg.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                             ctypes.c_void_p, ctypes.c_int]
        L.orc_transform.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_cell_boxes.restype = ctypes.c_int64
        L.orc_cell_boxes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_set_cell_matrices.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def regular_grid(d, level):
    """Reference slot order of AddRecursiveSonsOfPoint(center, level)."""
    L = lib()
    n = L.orc_regular_grid(d, level, None, None, 0)
    idx = np.zeros((n, d), dtype=np.int32)
    act = np.zeros(n, dtype=np.uint8)
    L.orc_regular_grid(d, level, _p(idx), _p(act), n)
    return idx, act


class Oracle:
    def __init__(self, idx, active=None):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        self.n, self.d = idx.shape
        self.active = (np.ones(self.n, dtype=np.uint8) if active is None
                       else np.ascontiguousarray(active, dtype=np.uint8))
        self._idx = idx
        self.h = lib().orc_create(self.d, self.n, _p(idx), _p(self.active))
        if not self.h:
            raise RuntimeError("orc_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_coef(self, kind, a0=1.0, c0=0.0, xi=None):
        xi = np.zeros(0) if xi is None else np.ascontiguousarray(xi, dtype=np.float64)
        lib().orc_set_coef(self.h, kind, a0, c0, _p(xi) if xi.size else None, int(xi.size))
        return self

    def cell_boxes(self):
        """(level, pos) of the oracle's cells of depth T+1 (ncells x d each)."""
        n = lib().orc_cell_boxes(self.h, None, None)
        lev = np.zeros((n, self.d), dtype=np.int32)
        pos = np.zeros((n, self.d), dtype=np.int32)
        lib().orc_cell_boxes(self.h, _p(lev), _p(pos))
        return lev, pos

    def set_cell_matrices(self, E):
        """Stored element matrices for the cells of cell_boxes() (None: back to the
        coefficient's element matrices)."""
        if E is None:
            self._mats = None
            lib().orc_set_cell_matrices(self.h, None)
        else:
            self._mats = np.ascontiguousarray(E, dtype=np.float64)
            lib().orc_set_cell_matrices(self.h, _p(self._mats))
        return self

    def depths(self):
        nd = lib().orc_num_depths(self.h)
        out = np.zeros((nd, self.d), dtype=np.int32)
        lib().orc_depths(self.h, _p(out))
        return out

    def multilevel_size(self):
        return lib().orc_multilevel_size(self.h)

    def sorted_depths(self, case):
        n = lib().orc_sorted_depths(self.h, case, None)
        out = np.zeros(max(n, 1), dtype=np.int32)
        lib().orc_sorted_depths(self.h, case, _p(out))
        return out[:n]

    def matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.n)
        lib().orc_matvec(self.h, _p(x), _p(y))
        return y

    def precond(self):
        z = np.zeros(self.n)
        lib().orc_precond(self.h, _p(z))
        return z

    def cg(self, f, eps=1e-10, maxit=500, x0=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        hist = np.zeros(maxit + 2)
        it = lib().orc_cg(self.h, eps, _p(f), _p(x), maxit, _p(hist), hist.size)
        return x, it, hist[: it + 1]

    # Basis transformations (complete grids): 0 hierarchize, 1 dehierarchize,
    # 2 prewavelets from nodal values, 3 nodal values from prewavelets.
    HIER, DEHIER, PREW_BY_NODAL, NODAL_BY_PREW = 0, 1, 2, 3

    def transform(self, what, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.n)
        rc = lib().orc_transform(self.h, what, _p(v), _p(out))
        if rc:
            raise ValueError("oracle transform: grid boxes are not complete (rc %d)" % rc)
        return out
