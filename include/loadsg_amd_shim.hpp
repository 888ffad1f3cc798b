/*
 * loadsg_amd_shim.hpp -- C++ drop-in for code written against LoAdSG's
 * MatrixVectorHomogen and CG::solveHomogen, running them on the GPU through the
 * C ABI of loadsg_amd.h.
 *
 * Include it after the reference's own headers (source/sparseEXPDE.h); link with
 * -lloadsg_amd.  It mirrors:
 *   MatrixVectorHomogen(AdaptiveSparseGrid&, MultiLevelAdaptiveSparseGrid&, bool symmetry, bool mpi_on)
 *       (MatrixVectorMultiplication/MatrixVectorHomogen.h:72)
 *   template<class P> void multiplication(VectorSparseG& prew, VectorSparseG& Ax, P& problem)
 *       (MatrixVectorHomogen.h:92,660-759): Ax is overwritten, as the reference does (Ax = 0.0 first)
 *   template<class S> static bool CG::solveHomogen(double eps, VectorSparseG& x, VectorSparseG& f,
 *       int* iterations, VectorSparseG& usol, MatrixVectorHomogen&, S stencil, double* time_precon)
 *       (applications/cg_method.h:27-33,759-886): x holds the start vector, the stopping rule is the
 *       reference's (iterate while r^T D^-1 r > eps, stop early once it is below eps^2) with the
 *       diagonal preconditioner precond3; the plug-in is
 *       taken by const reference (the reference copies it, stored matrices included).
 * Vectors stay the reference's VectorSparseG (host doubles in slot order); the device
 * copies live inside the operator.  The grid must be final before the operator is
 * built, as for VectorSparseG (vector.cc:12-23).
 *
 * Stencil plug-ins: Poisson and HelmHoltz map to LSG_COEF_CONST; a variable
 * coefficient is passed as lsg_shim::Coefficient (an lsg_coef, or host functions a(x),
 * c(x) sampled at the cell centres); any StencilTemplate plug-in goes through
 * lsg_shim::LocalStiffnessMatrices (its integration() evaluated once per cell and the
 * element matrices kept on the GPU, as the reference's LocalStiffnessMatrices keeps them on
 * its worker nodes).  Errors throw
 * std::runtime_error carrying lsg_last_error() (the reference prints and exits).
 */
#ifndef LOADSG_AMD_SHIM_HPP
#define LOADSG_AMD_SHIM_HPP

#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "loadsg_amd.h"

namespace lsg_shim {

inline void check(int rc) {
    if (rc != LSG_OK) throw std::runtime_error(std::string("loadsg_amd: ") + lsg_last_error());
}

/* A variable-coefficient plug-in: -div(a grad u) + c u with one of the LSG_COEF_* families,
 * or with general host functions a(x), c(x) (Coefficient::functions).  The latter is the
 * counterpart of a StencilTemplate subclass whose integration(cell, dirP, dirQ)
 * (stencils/Stencil.h:740) returns a(centre) * Poisson + c(centre) * HelmHoltz: the
 * library calls the functions at every cell centre when the operator is built. */
struct Coefficient {
    lsg_coef c;
    std::vector<double> xi;
    std::function<double(const double*)> fa, fc;   // LSG_COEF_FUNC
    std::string tag;    // names a function coefficient in the operator cache of MatrixVectorHomogen
    explicit Coefficient(int kind, double a0 = 1.0, double c0 = 0.0, std::vector<double> xi_ = {})
        : xi(std::move(xi_)) {
        std::memset(&c, 0, sizeof(c));
        c.kind = kind;
        c.a0 = a0;
        c.c0 = c0;
        bind();
    }
    /* Stored element matrices (LSG_COEF_MATRIX): entries stay owned by the caller until the
     * operator is built; tag names them in the operator cache. */
    static Coefficient matrices(std::string tag, const double* entries) {
        Coefficient k(LSG_COEF_MATRIX);
        k.c.cell_mat = entries;
        k.tag = std::move(tag);
        return k;
    }
    /* General coefficient from host functions of the point x (DimensionSparseGrid
     * coordinates); pass an empty std::function for an absent term.  Operators are cached
     * by tag: a different function needs a different tag. */
    static Coefficient functions(std::string tag, std::function<double(const double*)> a,
                                 std::function<double(const double*)> c) {
        Coefficient k(LSG_COEF_FUNC);
        k.fa = std::move(a);
        k.fc = std::move(c);
        k.tag = std::move(tag);
        k.bind();
        return k;
    }
    Coefficient(const Coefficient& o) : c(o.c), xi(o.xi), fa(o.fa), fc(o.fc), tag(o.tag) { bind(); }
    Coefficient& operator=(const Coefficient& o) {
        c = o.c;
        xi = o.xi;
        fa = o.fa;
        fc = o.fc;
        tag = o.tag;
        bind();
        return *this;
    }

  private:
    static double callA(const double* x, void* ctx) { return static_cast<Coefficient*>(ctx)->fa(x); }
    static double callC(const double* x, void* ctx) { return static_cast<Coefficient*>(ctx)->fc(x); }
    void bind() {   // c.xi and c.ctx always point at this object's own members
        c.nxi = (int)xi.size();
        c.xi = xi.empty() ? nullptr : xi.data();
        if (c.kind == LSG_COEF_FUNC) {
            c.a_fn = fa ? &Coefficient::callA : nullptr;
            c.c_fn = fc ? &Coefficient::callC : nullptr;
            c.ctx = this;
        }
    }
};

inline Coefficient coefficientOf(const Poisson&) { return Coefficient(LSG_COEF_CONST, 1.0, 0.0); }
inline Coefficient coefficientOf(const HelmHoltz&) { return Coefficient(LSG_COEF_CONST, 0.0, 1.0); }
inline Coefficient coefficientOf(const Coefficient& c) { return c; }

/* Device copy of the grid: the slot table of AdaptiveSparseGrid
 * (getIndexOfTable / getActiveTable, sgrid/sparseGrid.h:145-173). */
class DeviceGrid {
  public:
    explicit DeviceGrid(AdaptiveSparseGrid& grid) {
        const int d = DimensionSparseGrid;
        n_ = (int64_t)grid.getMaximalOccupiedSecondTable();
        std::vector<int32_t> idx((size_t)n_ * d);
        std::vector<uint8_t> act((size_t)n_);
        for (int64_t i = 0; i < n_; ++i) {
            IndexDimension I = grid.getIndexOfTable((unsigned long)i);
            for (int k = 0; k < d; ++k) idx[(size_t)i * d + k] = (int32_t)I.getIndex(k);
            act[(size_t)i] = grid.getActiveTable()[i] ? 1 : 0;
        }
        check(lsg_grid_create(d, n_, idx.data(), act.data(), &g_));
    }
    ~DeviceGrid() { lsg_grid_destroy(g_); }
    DeviceGrid(const DeviceGrid&) = delete;
    DeviceGrid& operator=(const DeviceGrid&) = delete;
    lsg_grid* get() const { return g_; }
    int64_t size() const { return n_; }

  private:
    lsg_grid* g_ = nullptr;
    int64_t n_ = 0;
};

/* LocalStiffnessMatrices (localStiffnessMatrices/LocalStiffnessMatrices.h:76-321) for the
 * device: the constructor evaluates the plug-in's integration(cell, dirP, dirQ)
 * (stencils/Stencil.h:740) once for every cell of depth T+1 of every depth T and every pair
 * of interior corners p <= q, after initialize(T) for the cell's depth, as
 * LocalStiffnessMatrixFixedDepthSymmetric::setValues does (LocalStiffnessMatrices.cpp:453-500);
 * addStencil adds a second plug-in's matrices (LocalStiffnessMatricesDynamicDistribution.h:
 * 181-197).  Passed to MatrixVectorHomogen::multiplication or CG::solveHomogen like any
 * plug-in, the matrices are uploaded once and applied on the GPU (the reference's
 * StoreLocalStiffnessMatrix branch, MatrixVectorHomogen.h:170-245,1619-1700; the
 * preconditioner is psi_p^T A psi_p from the same matrices, cg_method.h:544-641).
 * number_processes is accepted for the signature's sake: one GPU holds every depth. */
class LocalStiffnessMatrices {
  public:
    LocalStiffnessMatrices(AdaptiveSparseGrid& grid, StencilTemplate& stencil, int /*number_processes*/ = 1)
        : grid_(grid) {
        const int d = DimensionSparseGrid;
        check(lsg_grid_cells(grid_.get(), &ncells_, nullptr, nullptr));
        level_.resize((size_t)ncells_ * d);
        pos_.resize((size_t)ncells_ * d);
        check(lsg_grid_cell_boxes(grid_.get(), level_.data(), pos_.data()));
        entries_.assign((size_t)(ncells_ * lsg_cell_entries(d)), 0.0);
        add(stencil);
    }
    LocalStiffnessMatrices(const LocalStiffnessMatrices&) = delete;
    LocalStiffnessMatrices& operator=(const LocalStiffnessMatrices&) = delete;
    void addStencil(StencilTemplate& stencil) {
        add(stencil);
        ++version_;
    }
    TypeMatrixVectorMultiplication getTypeMatrixVectorMultiplication() { return StoreLocalStiffnessMatrix; }
    Coefficient coefficient() const {
        return Coefficient::matrices("lsm:" + std::to_string(reinterpret_cast<uintptr_t>(this)) + ":" +
                                         std::to_string(version_),
                                     entries_.data());
    }
    const std::vector<double>& entries() const { return entries_; }

  private:
    void add(StencilTemplate& stencil) {
        const int d = DimensionSparseGrid;
        const unsigned nc = 1u << d;
        const int64_t ne = lsg_cell_entries(d);
        Depth last;
        bool have = false;
        for (int64_t k = 0; k < ncells_; ++k) {
            const int32_t* lev = &level_[(size_t)k * d];
            const int32_t* pos = &pos_[(size_t)k * d];
            Depth T;
            for (int i = 0; i < d; ++i) T.set(lev[i], i);
            if (!have || !(T == last)) {
                stencil.initialize(T);
                last = T;
                have = true;
            }
            // the cell of depth T+1: 1D cell index 2^(T_i + 1) + pos_i (its centre
            // (2 pos_i + 1) 2^-(T_i + 1)), cells/celldimension.h:117-141
            CellDimension cell;
            for (int i = 0; i < d; ++i) cell.replace(i, (indexInteger)((1u << (lev[i] + 1)) + (unsigned)pos[i]));
            int e = 0;
            for (unsigned p = 0; p < nc; ++p)
                for (unsigned q = p; q < nc; ++q, ++e) {
                    bool interior = true;
                    for (int i = 0; i < d && interior; ++i) {
                        const int32_t cp = pos[i] + (int32_t)((p >> i) & 1u), cq = pos[i] + (int32_t)((q >> i) & 1u);
                        const int32_t top = int32_t(1) << lev[i];
                        interior = cp > 0 && cp < top && cq > 0 && cq < top;
                    }
                    if (!interior) continue;
                    CellIndexDirection dp, dq;
                    for (int i = 0; i < d; ++i) {
                        dp.setDirection(i, ((p >> i) & 1u) ? Right : Left);
                        dq.setDirection(i, ((q >> i) & 1u) ? Right : Left);
                    }
                    entries_[(size_t)(k * ne + e)] += stencil.integration(cell, dp, dq);
                }
        }
    }
    DeviceGrid grid_;
    int64_t ncells_ = 0;
    std::vector<int32_t> level_, pos_;
    std::vector<double> entries_;
    int version_ = 0;
};

inline Coefficient coefficientOf(const LocalStiffnessMatrices& m) { return m.coefficient(); }

/* MatrixVectorHomogen on the device.  One device operator per distinct coefficient
 * is created on first use and kept (the reference keeps its multilevel work arrays
 * in the MatrixVectorHomogen object in the same way). */
class MatrixVectorHomogen {
  public:
    MatrixVectorHomogen(AdaptiveSparseGrid& grid, MultiLevelAdaptiveSparseGrid& /*mgrid*/, bool /*symmetry*/ = true,
                        bool /*mpi_on*/ = false)
        : grid_(grid) {}
    ~MatrixVectorHomogen() {
        for (auto& e : ops_) lsg_op_destroy(e.op);
    }
    MatrixVectorHomogen(const MatrixVectorHomogen&) = delete;
    MatrixVectorHomogen& operator=(const MatrixVectorHomogen&) = delete;

    template <class P>
    void multiplication(VectorSparseG& prew, VectorSparseG& Ax, P& problem) {
        lsg_op* op = opFor(coefficientOf(problem));
        check(lsg_matvec(op, prew.getDatatableVector(), Ax.getDatatableVector(), LSG_HOST_SLOT));
    }

    lsg_op* opFor(const Coefficient& c) {
        for (auto& e : ops_)
            if (same(e.coef, c)) return e.op;
        lsg_op* op = nullptr;
        check(lsg_op_create(grid_.get(), &c.c, 64, 1, &op));
        ops_.push_back({coefficientOf(c), op});
        return op;
    }
    const DeviceGrid& deviceGrid() const { return grid_; }

  private:
    struct Entry {
        Coefficient coef;
        lsg_op* op;
    };
    static bool same(const Coefficient& a, const Coefficient& b) {
        if (a.c.kind != b.c.kind) return false;
        if (a.c.kind == LSG_COEF_FUNC || a.c.kind == LSG_COEF_MATRIX) return a.tag == b.tag;
        return a.c.a0 == b.c.a0 && a.c.c0 == b.c.c0 && a.xi == b.xi;
    }
    DeviceGrid grid_;
    std::vector<Entry> ops_;
};

/* CG::solveHomogen on the device (same arguments; usol is unused, as in the
 * reference).  *time_precon receives the seconds spent building the diagonal. */
struct CG {
    template <class S>
    static bool solveHomogen(double eps, VectorSparseG& x, VectorSparseG& f, int* iterations, VectorSparseG& /*usol*/,
                             MatrixVectorHomogen& matrix, const S& stencil, double* time_precon,
                             int maxit = 500 /* CG::maxIteration, cg_method.h:35 */) {
        lsg_op* op = matrix.opFor(coefficientOf(stencil));
        if (time_precon) {
            std::vector<double> diag((size_t)matrix.deviceGrid().size());
            const auto t0 = std::chrono::steady_clock::now();
            check(lsg_precond_diag(op, diag.data(), LSG_HOST_SLOT));
            *time_precon = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        lsg_solve_info info;
        std::memset(&info, 0, sizeof(info));
        const int rc = lsg_cg_solve(op, f.getDatatableVector(), x.getDatatableVector(), eps, LSG_STOP_REF_ABS, maxit,
                                    &info, LSG_HOST_SLOT);
        if (iterations) *iterations = info.iterations;
        if (rc != LSG_OK && rc != LSG_ERR_NOCONV) check(rc);
        return true;   // the reference always returns true
    }
};

/* Basis transformations on the device (BasisTransformations/BasisTransformations.cpp);
 * the reference's free functions take only the vectors, the shim also takes the
 * MatrixVectorHomogen that holds the device grid.  Grids with complete depth boxes only. */
inline void calcPrewByNodal(MatrixVectorHomogen& matrix, VectorSparseG& prew, VectorSparseG& u) {
    lsg_op* op = matrix.opFor(Coefficient(LSG_COEF_CONST, 1.0, 0.0));
    check(lsg_prew_by_nodal(op, u.getDatatableVector(), prew.getDatatableVector(), LSG_HOST_SLOT));
}
inline void calcNodalByPrew(MatrixVectorHomogen& matrix, VectorSparseG& prew, VectorSparseG& u) {
    lsg_op* op = matrix.opFor(Coefficient(LSG_COEF_CONST, 1.0, 0.0));
    check(lsg_nodal_by_prew(op, prew.getDatatableVector(), u.getDatatableVector(), LSG_HOST_SLOT));
}

/* H1::getValue / L2::getValue (applications/H1_Norm.cpp:8-78): the Poisson / mass energy
 * norm of the prewavelet coefficients of nodal values u. */
class H1 {
  public:
    H1(AdaptiveSparseGrid& grid, MultiLevelAdaptiveSparseGrid& mgrid) : m_(grid, mgrid) {}
    double getValue(VectorSparseG& u) { return norm(m_, u, Coefficient(LSG_COEF_CONST, 1.0, 0.0)); }
    static double norm(MatrixVectorHomogen& m, VectorSparseG& u, const Coefficient& c) {
        double out = 0.0;
        check(lsg_energy_norm(m.opFor(c), u.getDatatableVector(), &out, LSG_HOST_SLOT));
        return out;
    }

  private:
    MatrixVectorHomogen m_;
};
class L2 {
  public:
    L2(AdaptiveSparseGrid& grid, MultiLevelAdaptiveSparseGrid& mgrid) : m_(grid, mgrid) {}
    double getValue(VectorSparseG& u) { return H1::norm(m_, u, Coefficient(LSG_COEF_CONST, 0.0, 1.0)); }

  private:
    MatrixVectorHomogen m_;
};

}  // namespace lsg_shim

#endif /* LOADSG_AMD_SHIM_HPP */
