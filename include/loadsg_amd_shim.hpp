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
 *       reference's r^T D^-1 r < eps^2 with the diagonal preconditioner precond3.
 * Vectors stay the reference's VectorSparseG (host doubles in slot order); the device
 * copies live inside the operator.  The grid must be final before the operator is
 * built, as for VectorSparseG (vector.cc:12-23).
 *
 * Stencil plug-ins: Poisson and HelmHoltz map to LSG_COEF_CONST; a variable
 * coefficient is passed as lsg_shim::Coefficient (an lsg_coef, or host functions a(x),
 * c(x) sampled at the cell centres).  Errors throw
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
        if (a.c.kind == LSG_COEF_FUNC) return a.tag == b.tag;
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
                             MatrixVectorHomogen& matrix, S stencil, double* time_precon,
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
