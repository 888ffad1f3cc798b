"""loadsg_amd -- MI355X (gfx950) matrix-free prewavelet operator and CG solver
for sparse grids with zero Dirichlet boundaries (LoAdSG's hot path).

Thin ctypes layer over the C ABI in include/loadsg_amd.h.  The names mirror the
reference's C++ interface:

  AdaptiveSparseGrid            library/source/sgrid/sparseGrid.h:86-387
  MatrixVectorHomogen           library/source/MatrixVectorMultiplication/MatrixVectorHomogen.h:32-92
      .multiplication(x, Ax)    MatrixVectorHomogen.h:660-759
  Poisson / HelmHoltz           library/source/stencils/Stencil.h:873-988, 1605-1793
  VariableCoefficient           a StencilTemplate::integration plug-in (Stencil.h:740)
  CG.solveHomogen               library/source/applications/cg_method.h:27-33, 759-886
  Preconditioning.diagonal      library/source/applications/cg_method.h:82-451

Host vectors are float64 numpy arrays in the reference's slot order (the layout of
VectorSparseG's data table).  Device vectors are torch tensors on the GPU in this
library's depth-blocked order (see AdaptiveSparseGrid.device_order).

There is no CPU fallback: if the HIP library is not built, importing works but
every operation raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LOADSG_AMD_LIB") or os.path.join(_HERE, "_lib", "libloadsg_amd.so")

LSG_HOST_SLOT, LSG_DEVICE = 0, 1
LSG_ERR_NAN, LSG_ERR_NOCONV = -5, -6
STOP_REF_ABS, STOP_REL = 0, 1
CG_NO_PRECOND = 0x100     # or-ed into stop_rule: plain CG without the diagonal preconditioner
COEF_CONST, COEF_CFG2, COEF_CFG3, COEF_EXPSIN = 0, 1, 2, 3
COEF_CELL, COEF_FUNC = 4, 5
COEF_MATRIX, COEF_INTEGRATOR = 6, 7
PRECOND_DIAG, PRECOND_DEPTH_BLOCK = 0, 1

_lib = None


class LoadsgError(RuntimeError):
    pass


_COEF_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
_INTEG_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                             ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p)


class _Coef(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("a0", ctypes.c_double), ("c0", ctypes.c_double),
                ("nxi", ctypes.c_int), ("xi", ctypes.POINTER(ctypes.c_double)),
                ("cell_a", ctypes.POINTER(ctypes.c_double)), ("cell_c", ctypes.POINTER(ctypes.c_double)),
                ("a_fn", _COEF_FN), ("c_fn", _COEF_FN), ("ctx", ctypes.c_void_p),
                ("cell_mat", ctypes.POINTER(ctypes.c_double)), ("integ", _INTEG_FN)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("outer", ctypes.c_int), ("final_delta", ctypes.c_double),
                ("rel_residual", ctypes.c_double), ("converged", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def lib():
    """Load the gfx950 library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LoadsgError("libloadsg_amd.so not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `python loadsg_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
        sig = {
            "lsg_grid_create": (i32, [i32, i64, vp, vp, ctypes.POINTER(vp)]),
            "lsg_grid_regular": (i32, [i32, i32, ctypes.POINTER(vp)]),
            "lsg_grid_size": (i64, [vp]),
            "lsg_grid_dim": (i32, [vp]),
            "lsg_grid_slot_order": (i32, [vp, vp, vp]),
            "lsg_grid_device_order": (i32, [vp, vp]),
            "lsg_grid_stats": (i32, [vp, vp, i32]),
            "lsg_op_stats": (i32, [vp, vp, i32]),
            "lsg_grid_cells": (i32, [vp, ctypes.POINTER(i64), vp, vp]),
            "lsg_grid_cell_boxes": (i32, [vp, vp, vp]),
            "lsg_cell_entries": (i64, [i32]),
            "lsg_grid_destroy": (i32, [vp]),
            "lsg_op_create": (i32, [vp, ctypes.POINTER(_Coef), i32, i32, ctypes.POINTER(vp)]),
            "lsg_op_set_coef": (i32, [vp, i32, ctypes.POINTER(_Coef)]),
            "lsg_op_set_coefs": (i32, [vp, i32, i32, vp]),
            "lsg_op_batch": (i32, [vp]),
            "lsg_op_set_precond": (i32, [vp, i32]),
            "lsg_matvec": (i32, [vp, vp, vp, i32]),
            "lsg_precond_diag": (i32, [vp, vp, i32]),
            "lsg_cg_solve": (i32, [vp, vp, vp, dbl, i32, i32, vp, i32]),
            "lsg_ir_solve": (i32, [vp, vp, vp, vp, dbl, dbl, i32, vp, i32]),
            "lsg_ir_solve_many": (i32, [i32, vp, vp, vp, vp, dbl, dbl, i32, vp]),
            "lsg_rhs_sines": (i32, [vp, i32, vp, vp, vp, i32]),
            "lsg_rhs_nodes": (i32, [vp, ctypes.POINTER(i64), vp]),
            "lsg_rhs_separable": (i32, [vp, i32, vp, vp, vp, i32]),
            "lsg_last_timing": (i32, [vp, vp, i32]),
            "lsg_hierarchize": (i32, [vp, vp, vp, i32]),
            "lsg_dehierarchize": (i32, [vp, vp, vp, i32]),
            "lsg_prew_by_nodal": (i32, [vp, vp, vp, i32]),
            "lsg_nodal_by_prew": (i32, [vp, vp, vp, i32]),
            "lsg_energy_norm": (i32, [vp, vp, vp, i32]),
            "lsg_vector_norms": (i32, [vp, vp, vp, i32]),
            "lsg_op_set_block": (i32, [vp, i32, i32]),
            "lsg_op_destroy": (i32, [vp]),
            "lsg_mlp_create": (i32, [i32, i32, i32, i64, ctypes.c_uint64, ctypes.POINTER(vp)]),
            "lsg_mlp_forward": (i32, [vp, vp, vp, i32]),
            "lsg_mlp_gradients": (i32, [vp, vp, vp, i32, ctypes.POINTER(dbl)]),
            "lsg_mlp_train_step": (i32, [vp, vp, vp, i32, ctypes.c_float, ctypes.POINTER(dbl)]),
            "lsg_mlp_num_params": (i32, [vp, ctypes.POINTER(i64)]),
            "lsg_mlp_get_params": (i32, [vp, vp]),
            "lsg_mlp_set_params": (i32, [vp, vp]),
            "lsg_mlp_get_grads": (i32, [vp, vp]),
            "lsg_mlp_mix_rows": (i32, [vp, vp, i64, i32, vp, vp, i32, vp]),
            "lsg_mlp_destroy": (i32, [vp]),
            "lsg_gridgen_create": (vp, [i32]),
            "lsg_gridgen_destroy": (None, [vp]),
            "lsg_gridgen_add_point": (i32, [vp, vp]),
            "lsg_gridgen_add_recursive_sons": (i32, [vp, vp, i32]),
            "lsg_gridgen_refine_towards": (i32, [vp, dbl, i32, i32]),
            "lsg_gridgen_complete_dirichlet": (i32, [vp]),
            "lsg_gridgen_size": (i64, [vp]),
            "lsg_gridgen_points": (i32, [vp, vp, vp]),
            "lsg_last_error": (ctypes.c_char_p, []),
            "lsg_device_available": (i32, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise LoadsgError("loadsg_amd error %d: %s" % (rc, lib().lsg_last_error().decode()))


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)


def _sync_torch(*ts):
    for t in ts:
        if hasattr(t, "is_cuda") and t.is_cuda:
            import torch
            torch.cuda.synchronize(t.device)
            return


class GridGenerator:
    """The reference's grid generation calls on a growing point set (host code):
    AddPoint, AddRecursiveSonsOfPoint, completeDirichletGrid (sgrid/sparseGrid.cc:482-787),
    plus the refinement towards a corner of the locally adaptive benchmark grids.  Points
    are tuples of 1D indices (2 = 0.5, sons 2j and 2j+1); slots follow insertion order as
    in the reference's data table."""

    def __init__(self, d):
        self.d = d
        self._h = lib().lsg_gridgen_create(d)
        if not self._h:
            raise LoadsgError("bad dimension %d" % d)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lsg_gridgen_destroy(self._h)
            self._h = None

    def _gen(self, rc, what):
        if rc != 0:
            raise LoadsgError("loadsg_amd error %d: %s" % (rc, what))

    def AddPoint(self, index):
        v = np.ascontiguousarray(index, dtype=np.int32)
        self._gen(lib().lsg_gridgen_add_point(self._h, _ptr(v)), "AddPoint: boundary point or depth above 15")

    def AddRecursiveSonsOfPoint(self, index, level):
        v = np.ascontiguousarray(index, dtype=np.int32)
        self._gen(lib().lsg_gridgen_add_recursive_sons(self._h, _ptr(v), level),
                  "AddRecursiveSonsOfPoint: boundary point or depth above 15")

    def refine_towards(self, radius, max_level, corners=1):
        self._gen(lib().lsg_gridgen_refine_towards(self._h, radius, max_level, corners), "refine_towards: bad argument")

    def completeDirichletGrid(self):
        self._gen(lib().lsg_gridgen_complete_dirichlet(self._h), "completeDirichletGrid")

    @property
    def size(self):
        return lib().lsg_gridgen_size(self._h)

    def points(self):
        """(idx (size x d int32), active (size uint8)) in slot order."""
        idx = np.zeros((self.size, self.d), dtype=np.int32)
        act = np.zeros(self.size, dtype=np.uint8)
        self._gen(lib().lsg_gridgen_points(self._h, _ptr(idx), _ptr(act)), "points")
        return idx, act

    def grid(self):
        return AdaptiveSparseGrid(*self.points())

    @classmethod
    def refined(cls, d, start_level, radius, max_level, corners=1):
        """A regular start grid refined towards the corner(s), then completeDirichletGrid:
        the locally adaptive benchmark grids (start level 4, radius 0.15, d = 3)."""
        g = cls(d)
        g.AddRecursiveSonsOfPoint([2] * d, start_level)
        g.refine_towards(radius, max_level, corners)
        g.completeDirichletGrid()
        return g


class AdaptiveSparseGrid:
    """Sparse grid in the reference's slot order (AdaptiveSparseGrid, sparseGrid.h:86-387)."""

    def __init__(self, idx=None, active=None, _handle=None):
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            idx = np.ascontiguousarray(idx, dtype=np.int32)
            n, d = idx.shape
            act = None if active is None else np.ascontiguousarray(active, dtype=np.uint8)
            _check(lib().lsg_grid_create(d, n, _ptr(idx), _ptr(act), ctypes.byref(self._h)))

    @classmethod
    def regular(cls, d, level):
        """AddRecursiveSonsOfPoint(center, level) (sparseGrid.cc:557-566)."""
        h = ctypes.c_void_p()
        _check(lib().lsg_grid_regular(d, level, ctypes.byref(h)))
        return cls(_handle=h)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.lsg_grid_destroy(self._h)
            self._h = ctypes.c_void_p()

    @property
    def dim(self):
        return lib().lsg_grid_dim(self._h)

    def getMaximalOccupiedSecondTable(self):
        return lib().lsg_grid_size(self._h)

    size = property(getMaximalOccupiedSecondTable)

    def slot_order(self):
        n, d = self.size, self.dim
        idx = np.zeros((n, d), dtype=np.int32)
        act = np.zeros(n, dtype=np.uint8)
        _check(lib().lsg_grid_slot_order(self._h, _ptr(idx), _ptr(act)))
        return idx, act

    def device_order(self):
        """dev_of_slot[i] = device position of reference slot i."""
        out = np.zeros(self.size, dtype=np.int64)
        _check(lib().lsg_grid_device_order(self._h, _ptr(out)))
        return out

    def cells(self):
        """Cells of depth T+1 of every depth T (CellData, cells/CellStructure.h:23-61), in the
        order a CellCoefficient is read: (centres (ncells x d), depth-list index of each cell)."""
        n = ctypes.c_int64()
        _check(lib().lsg_grid_cells(self._h, ctypes.byref(n), None, None))
        x = np.zeros((n.value, self.dim))
        t = np.zeros(n.value, dtype=np.int32)
        _check(lib().lsg_grid_cells(self._h, ctypes.byref(n), _ptr(x), _ptr(t)))
        return x, t

    def cell_boxes(self):
        """The same cells as level-T boxes: (level, pos), both ncells x d int32; cell k is
        prod_i [pos_i 2^-level_i, (pos_i + 1) 2^-level_i]."""
        n = self.cells()[0].shape[0]
        lev = np.zeros((n, self.dim), dtype=np.int32)
        pos = np.zeros((n, self.dim), dtype=np.int32)
        _check(lib().lsg_grid_cell_boxes(self._h, _ptr(lev), _ptr(pos)))
        return lev, pos

    def stats(self):
        out = np.zeros(12, dtype=np.int64)
        _check(lib().lsg_grid_stats(self._h, _ptr(out), 12))
        keys = ["depths", "multilevel_size", "apply_items", "wavefront_levels", "stream_values", "max_box",
                "case_buffer", "w_buffer", "part_buffer", "moved_values", "apply_values", "stream_apply_values"]
        return dict(zip(keys, (int(v) for v in out)))


def _coef(kind, a0=1.0, c0=0.0, xi=None):
    c = _Coef()
    c.kind, c.a0, c.c0 = kind, a0, c0
    if xi is not None and len(xi):
        arr = (ctypes.c_double * len(xi))(*[float(v) for v in xi])
        c.nxi, c.xi = len(xi), ctypes.cast(arr, ctypes.POINTER(ctypes.c_double))
        c._keep = arr
    else:
        c.nxi, c.xi = 0, None
    return c


class Stencil:
    """Operator plug-in: -div(a grad u) + c u with the coefficient family `kind`."""

    def __init__(self, kind, a0=1.0, c0=0.0, xi=None):
        self.kind, self.a0, self.c0, self.xi = kind, a0, c0, (None if xi is None else list(xi))

    def _c(self, grid):
        return _coef(self.kind, self.a0, self.c0, self.xi)


def _dbl_array(v, n, name):
    a = np.ascontiguousarray(v, dtype=np.float64)
    if a.shape != (n,):
        raise LoadsgError("%s: expected %d cell values, got shape %s" % (name, n, a.shape))
    return a


class CellCoefficient(Stencil):
    """General coefficient given per cell: a[k], c[k] at the centre of cell k of
    AdaptiveSparseGrid.cells() (None: no stiffness / no mass term).  The element matrix of
    cell k is a[k] * Poisson + c[k] * HelmHoltz (a StencilTemplate::integration plug-in,
    Stencil.h:740)."""

    def __init__(self, a=None, c=None):
        if a is None and c is None:
            raise LoadsgError("CellCoefficient needs a or c")
        self.kind, self.a, self.c = COEF_CELL, a, c

    def _c(self, grid):
        n = grid.cells()[0].shape[0]
        cc = _Coef()
        cc.kind = COEF_CELL
        keep = []
        for field, v in (("cell_a", self.a), ("cell_c", self.c)):
            if v is not None:
                arr = _dbl_array(v, n, field)
                keep.append(arr)
                setattr(cc, field, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        cc._keep = keep
        return cc


class CellFunction(Stencil):
    """General coefficient from vectorised functions of the cell centres: a(x), c(x) take
    an (ncells x d) array and return ncells values; sampled through grid.cells()."""

    def __init__(self, a=None, c=None):
        if a is None and c is None:
            raise LoadsgError("CellFunction needs a or c")
        self.kind, self.a, self.c = COEF_CELL, a, c

    def _c(self, grid):
        x, _ = grid.cells()
        return CellCoefficient(None if self.a is None else self.a(x), None if self.c is None else self.c(x))._c(grid)


class HostFunction(Stencil):
    """General coefficient from scalar host functions a(x), c(x) of one point (a length-d
    array), called by the library at every cell centre (LSG_COEF_FUNC)."""

    def __init__(self, a=None, c=None):
        if a is None and c is None:
            raise LoadsgError("HostFunction needs a or c")
        self.kind, self.a, self.c = COEF_FUNC, a, c

    def _c(self, grid):
        d = grid.dim
        cc = _Coef()
        cc.kind = COEF_FUNC
        keep = []
        for field, fn in (("a_fn", self.a), ("c_fn", self.c)):
            if fn is not None:
                cb = _COEF_FN(lambda x, ctx, fn=fn: float(fn(np.ctypeslib.as_array(x, (d,)))))
                keep.append(cb)
                setattr(cc, field, cb)
        cc._keep = keep
        return cc


def cell_entries(d):
    """Values of one stored element matrix: 2^(d-1) (2^d + 1) (corners p <= q)."""
    return (1 << (d - 1)) * ((1 << d) + 1)


def cell_pairs(d):
    """Corner pairs (p, q), p <= q, in stored-entry order (the reference's nested
    CellIndexIterator loops, cells/iterator_neu.cpp:34-49); bit i of a corner = upper end of
    the cell in dimension i."""
    nc = 1 << d
    return [(p, q) for p in range(nc) for q in range(p, nc)]


class LocalStiffnessMatrices(Stencil):
    """Stored element matrices (LocalStiffnessMatrices, localStiffnessMatrices/
    LocalStiffnessMatrices.h:76-321): one symmetric 2^d x 2^d matrix per cell of
    grid.cells(), computed once from a plug-in's integration and applied by the matvec in
    place of the on-the-fly stencil (MatrixVectorHomogen.h:170-245,1619-1700).

    integration: an (ncells x cell_entries(d)) array in cell_pairs(d) order, or a function
    integration(level, pos, p, q) of arrays of cell boxes (ncells x d each, as
    grid.cell_boxes()) and one corner pair, returning ncells values."""

    def __init__(self, grid, integration):
        self.kind = COEF_MATRIX
        d = grid.dim
        ncell = grid.cells()[0].shape[0]
        if callable(integration):
            lev, pos = grid.cell_boxes()
            E = np.zeros((ncell, cell_entries(d)))
            for e, (p, q) in enumerate(cell_pairs(d)):
                E[:, e] = integration(lev, pos, p, q)
        else:
            E = np.array(integration, dtype=np.float64)
        if E.shape != (ncell, cell_entries(d)):
            raise LoadsgError("element matrices: expected shape (%d, %d), got %s" % (ncell, cell_entries(d), E.shape))
        self.entries = np.ascontiguousarray(E)
        self.ncell = ncell

    def addStencil(self, other):
        """Adds the element matrices of another LocalStiffnessMatrices on the same grid
        (LocalStiffnessMatricesDynamicDistribution::addStencil, ADDLocalStiffnessMatrices)."""
        if other.entries.shape != self.entries.shape:
            raise LoadsgError("addStencil: element matrices of another grid")
        self.entries = self.entries + other.entries
        return self

    def _c(self, grid):
        if grid.cells()[0].shape[0] != self.ncell:
            raise LoadsgError("element matrices were built for another grid")
        cc = _Coef()
        cc.kind = COEF_MATRIX
        cc.cell_mat = self.entries.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        cc._keep = [self.entries]
        return cc


class CellIntegrator(Stencil):
    """Stored element matrices from a scalar host integrator integration(level, pos, p, q)
    of one cell (level, pos: length-d int arrays) and one corner pair, called by the library
    once per cell and interior pair (LSG_COEF_INTEGRATOR: a plug-in's
    StencilTemplate::integration(cell, dirP, dirQ), Stencil.h:740)."""

    def __init__(self, integration):
        self.kind, self.fn = COEF_INTEGRATOR, integration

    def _c(self, grid):
        fn = self.fn
        cb = _INTEG_FN(lambda d, lev, pos, p, q, ctx: float(
            fn(np.ctypeslib.as_array(lev, (d,)), np.ctypeslib.as_array(pos, (d,)), int(p), int(q))))
        cc = _Coef()
        cc.kind = COEF_INTEGRATOR
        cc.integ = cb
        cc._keep = [cb]
        return cc


def Poisson():
    """Poisson (Stencil.h:1605-1793)."""
    return Stencil(COEF_CONST, 1.0, 0.0)


def HelmHoltz():
    """HelmHoltz = Q1 mass (Stencil.h:873-988)."""
    return Stencil(COEF_CONST, 0.0, 1.0)


def VariableCoefficient(kind, xi=None):
    return Stencil(kind, xi=xi)


class MatrixVectorHomogen:
    """Matrix-free prewavelet operator (MatrixVectorHomogen.h:32-92) on one grid.

    precision 64 or 32; batch instances share the grid (coefficients per instance
    via set_coefficient)."""

    def __init__(self, grid, stencil, precision=64, batch=1):
        self.grid, self.precision, self.batch = grid, precision, batch
        self._h = ctypes.c_void_p()
        c = stencil._c(grid)
        _check(lib().lsg_op_create(grid._h, ctypes.byref(c), precision, batch, ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.lsg_op_destroy(self._h)
            self._h = ctypes.c_void_p()

    def set_coefficient(self, instance, stencil):
        c = stencil._c(self.grid)
        _check(lib().lsg_op_set_coef(self._h, instance, ctypes.byref(c)))

    def set_coefficients(self, stencils, first=0):
        """set_coefficient for instances first, first + 1, ... in one library call."""
        cs = [st._c(self.grid) for st in stencils]
        arr = (_Coef * len(cs))(*cs)
        _check(lib().lsg_op_set_coefs(self._h, first, len(cs), ctypes.cast(arr, ctypes.c_void_p)))

    def set_block(self, threads):
        _check(lib().lsg_op_set_block(self._h, 0, threads))

    def set_precond(self, kind):
        """Preconditioner of CG.solveHomogen and of the inner solves of CG.solveMixed on this
        operator: PRECOND_DIAG (the reference's precond3, default) or PRECOND_DEPTH_BLOCK
        (block Jacobi over the depth blocks with constant-coefficient Galerkin blocks scaled
        by the diagonal; on adaptive grids over the active points of each depth)."""
        _check(lib().lsg_op_set_precond(self._h, kind))
        return self

    # ---- argument checks: the C ABI takes raw pointers to batch x N values
    def _host(self, v, name="vector", out=False):
        """float64 C-contiguous numpy array of batch x N values in slot order (shape (N,)
        allowed for batch 1); out=True also requires a writable array of that dtype."""
        n = self.grid.size
        if out:
            if not (isinstance(v, np.ndarray) and v.dtype == np.float64 and v.flags.c_contiguous and v.flags.writeable):
                raise LoadsgError("%s: expected a writable C-contiguous float64 numpy array" % name)
            a = v
        else:
            a = np.ascontiguousarray(v, dtype=np.float64)
        if a.size != self.batch * n or (self.batch > 1 and a.shape != (self.batch, n)) or a.ndim > 2:
            raise LoadsgError("%s: expected shape (%d, %d)%s, got %s" % (
                name, self.batch, n, " or (%d,)" % n if self.batch == 1 else "", a.shape))
        return a

    def _dev(self, t, name="tensor", fp64=None):
        """Contiguous GPU tensor of batch x N values in device order, of the operator's
        precision (fp64=True/False overrides it)."""
        import torch
        want = torch.float64 if (self.precision == 64 if fp64 is None else fp64) else torch.float32
        if not (t.is_cuda and t.dtype == want and t.is_contiguous() and t.numel() == self.batch * self.grid.size):
            raise LoadsgError("%s: expected a contiguous %s GPU tensor of %d x %d values, got %s %s %s" % (
                name, want, self.batch, self.grid.size, t.dtype, tuple(t.shape), t.device))
        return t

    def _new_host(self):
        return np.zeros((self.batch, self.grid.size)) if self.batch > 1 else np.zeros(self.grid.size)

    def multiplication(self, prew, Ax=None):
        """Ax = A prew.  numpy float64 (batch x N, or N for batch 1; slot order) or torch
        device tensors (device order)."""
        if hasattr(prew, "data_ptr"):
            self._dev(prew, "prew")
            self._dev(Ax, "Ax")
            _sync_torch(prew, Ax)
            _check(lib().lsg_matvec(self._h, _ptr(prew), _ptr(Ax), LSG_DEVICE))
            return Ax
        x = self._host(prew, "prew")
        y = np.zeros_like(x) if Ax is None else self._host(Ax, "Ax", out=True)
        _check(lib().lsg_matvec(self._h, _ptr(x), _ptr(y), LSG_HOST_SLOT))
        return y

    def diagonal(self, out=None):
        """Preconditioning<S>::precond3 diagonal (cg_method.h:322-451)."""
        if out is not None and hasattr(out, "data_ptr"):
            self._dev(out, "out")
            _sync_torch(out)
            _check(lib().lsg_precond_diag(self._h, _ptr(out), LSG_DEVICE))
            return out
        z = self._new_host()
        _check(lib().lsg_precond_diag(self._h, _ptr(z), LSG_HOST_SLOT))
        return z

    def rhs_sines(self, eta, m, out=None):
        """b_p = int f psi_p for f = sum_j eta_j prod_i sin(m_j pi x_i); eta: batch x nterms."""
        eta = np.ascontiguousarray(np.atleast_2d(eta), dtype=np.float64)
        m = np.ascontiguousarray(m, dtype=np.int32)
        if eta.shape != (self.batch, m.size):
            raise LoadsgError("eta: expected shape (%d, %d), got %s" % (self.batch, m.size, eta.shape))
        if out is not None and hasattr(out, "data_ptr"):
            self._dev(out, "out")
            _sync_torch(out)
            _check(lib().lsg_rhs_sines(self._h, m.size, _ptr(eta), _ptr(m), _ptr(out), LSG_DEVICE))
            return out
        b = self._new_host()
        _check(lib().lsg_rhs_sines(self._h, m.size, _ptr(eta), _ptr(m), _ptr(b), LSG_HOST_SLOT))
        return b

    def rhs_nodes(self):
        """1D quadrature nodes at which rhs_separable reads its factors."""
        n = ctypes.c_int64()
        _check(lib().lsg_rhs_nodes(self._h, ctypes.byref(n), None))
        x = np.zeros(n.value)
        _check(lib().lsg_rhs_nodes(self._h, ctypes.byref(n), _ptr(x)))
        return x

    def rhs_separable(self, factors, weights=None, out=None):
        """Load vector b_p = int f psi_p of f = sum_j w_j prod_i g_ji(x_i).  factors[j][i] is a
        vectorised function of one coordinate (an array of points -> values); weights has
        shape (batch, nterms) (default: all ones)."""
        d = self.grid.dim
        x = self.rhs_nodes()
        nt = len(factors)
        vals = np.zeros((nt, d, x.size))
        for j, fac in enumerate(factors):
            if len(fac) != d:
                raise LoadsgError("term %d: expected %d factors" % (j, d))
            for i, g in enumerate(fac):
                vals[j, i] = np.broadcast_to(np.asarray(g(x), dtype=np.float64), x.shape)
        w = np.ones((self.batch, nt)) if weights is None else np.ascontiguousarray(np.atleast_2d(weights),
                                                                                   dtype=np.float64)
        if w.shape != (self.batch, nt):
            raise LoadsgError("weights: expected shape (%d, %d), got %s" % (self.batch, nt, w.shape))
        vals = np.ascontiguousarray(vals)
        if out is not None and hasattr(out, "data_ptr"):
            self._dev(out, "out")
            _sync_torch(out)
            _check(lib().lsg_rhs_separable(self._h, nt, _ptr(vals), _ptr(w), _ptr(out), LSG_DEVICE))
            return out
        b = self._new_host()
        _check(lib().lsg_rhs_separable(self._h, nt, _ptr(vals), _ptr(w), _ptr(b), LSG_HOST_SLOT))
        return b

    # ---- basis transformations (FP64 operators, grids with complete depth boxes)
    def _transform(self, fn, v, out):
        if hasattr(v, "data_ptr"):
            self._dev(v, "input")
            self._dev(out, "out")
            _sync_torch(v, out)
            _check(fn(self._h, _ptr(v), _ptr(out), LSG_DEVICE))
            return out
        x = self._host(v, "input")
        y = np.zeros_like(x) if out is None else self._host(out, "out", out=True)
        _check(fn(self._h, _ptr(x), _ptr(y), LSG_HOST_SLOT))
        return y

    def hierarchize(self, u, out=None):
        """CalcHierarchicalBasis (BasisTransformations.cpp:390-409)."""
        return self._transform(lib().lsg_hierarchize, u, out)

    def dehierarchize(self, hier, out=None):
        """CalcNodalBasis (BasisTransformations.cpp:174-198)."""
        return self._transform(lib().lsg_dehierarchize, hier, out)

    def prew_by_nodal(self, u, out=None):
        """calcPrewByNodal (BasisTransformations.cpp:486-545)."""
        return self._transform(lib().lsg_prew_by_nodal, u, out)

    def nodal_by_prew(self, prew, out=None):
        """calcNodalByPrew (BasisTransformations.cpp:670-707)."""
        return self._transform(lib().lsg_nodal_by_prew, prew, out)

    def energy_norm(self, u):
        """sqrt(v^T A v), v = prewavelets of nodal values u, per instance
        (H1::getValue / L2::getValue, applications/H1_Norm.cpp:8-78)."""
        out = np.zeros(self.batch)
        if hasattr(u, "data_ptr"):
            self._dev(u, "u")
            _sync_torch(u)
            _check(lib().lsg_energy_norm(self._h, _ptr(u), _ptr(out), LSG_DEVICE))
        else:
            x = self._host(u, "u")
            _check(lib().lsg_energy_norm(self._h, _ptr(x), _ptr(out), LSG_HOST_SLOT))
        return out if self.batch > 1 else float(out[0])

    def vector_norms(self, x):
        """(L_infty, L_2) of x per instance (applications/norms.h:12-100): max |x| over all
        points and sqrt(sum x^2 / dofs) over the active ones."""
        out = np.zeros((self.batch, 2))
        if hasattr(x, "data_ptr"):
            self._dev(x, "x")
            _sync_torch(x)
            _check(lib().lsg_vector_norms(self._h, _ptr(x), _ptr(out), LSG_DEVICE))
        else:
            v = self._host(x, "x")
            _check(lib().lsg_vector_norms(self._h, _ptr(v), _ptr(out), LSG_HOST_SLOT))
        return out if self.batch > 1 else out[0]

    def stats(self):
        """Statistics of the plan this operator runs on (keys as AdaptiveSparseGrid.stats)."""
        out = np.zeros(12, dtype=np.int64)
        _check(lib().lsg_op_stats(self._h, _ptr(out), 12))
        keys = ["depths", "multilevel_size", "apply_items", "wavefront_levels", "stream_values", "max_box",
                "case_buffer", "w_buffer", "part_buffer", "moved_values", "apply_values", "stream_apply_values"]
        return dict(zip(keys, (int(v) for v in out)))

    def last_timing(self):
        out = np.zeros(3)
        _check(lib().lsg_last_timing(self._h, _ptr(out), 3))
        return {"total_ms": out[0], "apply_ms": out[1], "apply_launches": int(out[2])}


def CalcHierarchicalBasis(u, op):
    """Hierarchical surpluses of nodal values u (BasisTransformations.cpp:390-409); op is
    an FP64 MatrixVectorHomogen on u's grid (it holds the device context)."""
    return op.hierarchize(u)


def CalcNodalBasis(hier, op):
    """Nodal values from hierarchical surpluses (BasisTransformations.cpp:174-198)."""
    return op.dehierarchize(hier)


def calcPrewByNodal(u, op):
    """Prewavelet coefficients of nodal values u (BasisTransformations.cpp:486-545)."""
    return op.prew_by_nodal(u)


def calcNodalByPrew(prew, op):
    """Nodal values of a prewavelet expansion (BasisTransformations.cpp:670-707)."""
    return op.nodal_by_prew(prew)


def L_infty(x, op):
    """max |x_i| over the stored points (applications/norms.h:12-52), on the device."""
    return op.vector_norms(x)[..., 0]


def L_2(x, op):
    """sqrt(sum x_i^2 / dofs) over the active points (applications/norms.h:54-100)."""
    return op.vector_norms(x)[..., 1]


class H1:
    """H1 energy norm of nodal values: H1::getValue (applications/H1_Norm.cpp:8-41)."""

    def __init__(self, grid, batch=1):
        self.op = MatrixVectorHomogen(grid, Poisson(), 64, batch)

    def getValue(self, u):
        return self.op.energy_norm(u)


class L2:
    """L2 norm of the sparse-grid interpolant of nodal values: L2::getValue
    (applications/H1_Norm.cpp:45-78), the mass-matrix energy norm."""

    def __init__(self, grid, batch=1):
        self.op = MatrixVectorHomogen(grid, HelmHoltz(), 64, batch)

    def getValue(self, u):
        return self.op.energy_norm(u)


class CG:
    """CG::solveHomogen (cg_method.h:759-886) on the device."""

    @staticmethod
    def solveHomogen(eps, x, f, matrix, stop_rule=STOP_REF_ABS, maxit=500):
        """Returns the solve info (a list for batches); an instance that stops at maxit has
        converged == 0 (the library's status LSG_ERR_NOCONV is not raised, as the reference
        returns normally after CG::maxIteration, cg_method.h:852)."""
        info = (SolveInfo * matrix.batch)()
        if hasattr(x, "data_ptr"):
            # x is FP64 on FP64 operators and of the operator's precision otherwise
            matrix._dev(x, "x")
            matrix._dev(f, "f")
            _sync_torch(x, f)
            rc = lib().lsg_cg_solve(matrix._h, _ptr(f), _ptr(x), eps, stop_rule, maxit, info, LSG_DEVICE)
        else:
            xa = matrix._host(x, "x")
            fa = matrix._host(f, "f")
            rc = lib().lsg_cg_solve(matrix._h, _ptr(fa), _ptr(xa), eps, stop_rule, maxit, info, LSG_HOST_SLOT)
            x[...] = xa.reshape(np.shape(x))
        if rc != LSG_ERR_NOCONV:
            _check(rc)
        res = [info[b].as_dict() for b in range(matrix.batch)]
        return res[0] if matrix.batch == 1 else res

    @staticmethod
    def solveMixed(rel_tol, x, f, op64, op32, inner_rel_tol=1e-4, maxit=30):
        """FP32 inner PCG under FP64 iterative refinement (all reductions FP64)."""
        info = (SolveInfo * op64.batch)()
        if hasattr(x, "data_ptr"):
            op64._dev(x, "x")
            op64._dev(f, "f")
            _sync_torch(x, f)
            rc = lib().lsg_ir_solve(op64._h, op32._h, _ptr(f), _ptr(x), rel_tol, inner_rel_tol, maxit, info,
                                    LSG_DEVICE)
        else:
            xa = op64._host(x, "x")
            fa = op64._host(f, "f")
            rc = lib().lsg_ir_solve(op64._h, op32._h, _ptr(fa), _ptr(xa), rel_tol, inner_rel_tol, maxit, info,
                                    LSG_HOST_SLOT)
            x[...] = xa.reshape(np.shape(x))
        if rc != LSG_ERR_NOCONV:
            _check(rc)
        res = [info[b].as_dict() for b in range(op64.batch)]
        return res[0] if op64.batch == 1 else res


def solve_many(problems, rel_tol, inner_rel_tol=1e-4, maxit=30):
    """Mixed-precision solves of instances on different grids at once (lsg_ir_solve_many).

    problems: list of (x, f, op64, op32), each an operator pair with its own grid and
    batch and device tensors x, f of shape (batch, N) (see CG.solveMixed).  One host
    thread advances all groups in lock-step; the GPU runs them side by side on the
    operators' streams.  Returns the per-group lists of solve infos."""
    G = len(problems)
    if G == 0:
        return []
    ops64 = (ctypes.c_void_p * G)()
    ops32 = (ctypes.c_void_p * G)()
    fs = (ctypes.c_void_p * G)()
    xs = (ctypes.c_void_p * G)()
    infos = [(SolveInfo * p[2].batch)() for p in problems]
    iptr = (ctypes.c_void_p * G)()
    for q, (x, f, o64, o32) in enumerate(problems):
        o64._dev(x, "x")
        o64._dev(f, "f")
        ops64[q], ops32[q] = o64._h.value, o32._h.value
        fs[q], xs[q] = f.data_ptr(), x.data_ptr()
        iptr[q] = ctypes.cast(infos[q], ctypes.c_void_p).value
    _sync_torch(*[p[0] for p in problems])
    rc = lib().lsg_ir_solve_many(G, ops64, ops32, fs, xs, rel_tol, inner_rel_tol, maxit, iptr)
    if rc != LSG_ERR_NOCONV:
        _check(rc)
    return [[inf[b].as_dict() for b in range(p[2].batch)] for inf, p in zip(infos, problems)]


def device_available():
    return bool(lib().lsg_device_available())


class MLP:
    """Warm-start network n_in -> width (x hidden, GELU) -> n_out on the GPU (BF16
    matrix-core GEMMs, FP32 accumulation and master weights, Adam).  x / target /
    output are float32 torch tensors on the GPU, shape (batch, n_in) / (batch, n_out)."""

    def __init__(self, n_in, width, hidden, n_out, seed=0):
        self._h = ctypes.c_void_p()
        self.n_in, self.width, self.hidden, self.n_out = n_in, width, hidden, n_out
        _check(lib().lsg_mlp_create(n_in, width, hidden, n_out, seed, ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lsg_mlp_destroy(self._h)
            self._h = None

    @staticmethod
    def _need(t, rows, cols):
        import torch
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == 2 and
                t.shape[1] == cols and (rows is None or t.shape[0] == rows)):
            raise LoadsgError("expected a contiguous float32 GPU tensor of shape (batch, %d)" % cols)

    def forward(self, x, out=None):
        import torch
        self._need(x, None, self.n_in)
        y = torch.empty(x.shape[0], self.n_out, dtype=torch.float32, device=x.device) if out is None else out
        self._need(y, x.shape[0], self.n_out)
        _sync_torch(x)
        _check(lib().lsg_mlp_forward(self._h, _ptr(x), _ptr(y), x.shape[0]))
        return y

    def gradients(self, x, target):
        self._need(x, None, self.n_in)
        self._need(target, x.shape[0], self.n_out)
        _sync_torch(x)
        loss = ctypes.c_double()
        _check(lib().lsg_mlp_gradients(self._h, _ptr(x), _ptr(target), x.shape[0], ctypes.byref(loss)))
        return loss.value

    def train_step(self, x, target, lr):
        self._need(x, None, self.n_in)
        self._need(target, x.shape[0], self.n_out)
        _sync_torch(x)
        loss = ctypes.c_double()
        _check(lib().lsg_mlp_train_step(self._h, _ptr(x), _ptr(target), x.shape[0], lr, ctypes.byref(loss)))
        return loss.value

    def mix_rows(self, U, rows, w, out=None):
        """out[b] = sum_j w[b, j] U[rows[b], j] for U (S x J x n), rows (batch,) int64 and
        w (batch x J), all FP32 / int64 device tensors."""
        import torch
        S, J, n = U.shape
        assert U.dtype == torch.float32 and U.is_contiguous() and rows.dtype == torch.int64
        assert w.dtype == torch.float32 and w.is_contiguous() and tuple(w.shape) == (rows.shape[0], J)
        out = torch.empty(rows.shape[0], n, dtype=torch.float32, device=U.device) if out is None else out
        _sync_torch(U)
        _check(lib().lsg_mlp_mix_rows(self._h, _ptr(U), n, J, _ptr(rows), _ptr(w), rows.shape[0], _ptr(out)))
        return out

    def num_params(self):
        n = ctypes.c_int64()
        _check(lib().lsg_mlp_num_params(self._h, ctypes.byref(n)))
        return n.value

    def params(self):
        """Flat float32 copy: per layer W (in x out, row-major) then b."""
        a = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().lsg_mlp_get_params(self._h, _ptr(a)))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.num_params():
            raise LoadsgError("parameter count mismatch")
        _check(lib().lsg_mlp_set_params(self._h, _ptr(a)))

    def grads(self):
        a = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().lsg_mlp_get_grads(self._h, _ptr(a)))
        return a

    def layer_shapes(self):
        dims = [self.n_in] + [self.width] * self.hidden + [self.n_out]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]
