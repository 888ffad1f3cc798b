// lsg_kernels.hpp -- launch interface of the gfx950 kernels (lsg_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "lsg_plan.hpp"

namespace lsg {

constexpr int MAXG = 64;   // most batch instances one apply workgroup handles
constexpr int WSHAPE_BYTES = 176;   // per-wave box-shape block in LDS (apply kernel)

// Per-instance CG state, device resident (scalars of CG::solveHomogen,
// applications/cg_method.h:763-883).
struct CGState {
    double delta, rr, ff, tau, beta, eps;
    int it, active, do_dir, converged, maxit, rule;
};

template <typename V>
struct ApplyArgs {
    const FamilyItem* fam;    // families of one wavefront level
    const ApplyItem* items;   // all apply items
    int d;
    int general;          // 0: one separable stiffness term (2 LDS boxes), 1: general sum of terms
    int nA, nC;           // stiffness / mass terms
    const V* coefTab;     // [inst][term][dim][tabLen]
    int64_t coefStride;
    int64_t tabLen;
    const V* nodal;
    int64_t nStride;
    const V* cbuf;
    int64_t cStride;
    V* wbuf;
    int64_t wStride;
    V* part;
    int64_t pStride;
    int64_t maxBox;       // LDS box size of this launch
    int64_t maxW;         // LDS weight-table size of this launch
    int G;                // batch instances per workgroup
    int cbits;            // bits per packed point coordinate (0: recompute coordinates per sweep)
    int batch;
    const CGState* st;    // may be null
};

template <typename V>
void launch_nodal(const NodalItem* items, int nitems, int d, const V* x, int64_t xStride, V* nodal, int64_t nStride,
                  int64_t maxBox, int batch, const CGState* st, hipStream_t s);
template <typename V>
void launch_prol(const ProlItem* items, int nitems, int d, const V* nodal, int64_t nStride, V* cbuf, int64_t cStride,
                 int64_t maxBox, int batch, const CGState* st, hipStream_t s);
template <typename V>
int launch_apply(const ApplyArgs<V>& a, int nitems, int batch, int threads, size_t lds, hipStream_t s);
template <typename V>
void launch_sum(const SumItem* items, int nitems, const int64_t* list, int d, const V* part, int64_t pStride, V* y,
                int64_t yStride, int batch, const CGState* st, hipStream_t s);
template <typename V>
void launch_diag(const SumItem* items, int nitems, int d, int nA, int nC, const V* coefTab, int64_t coefStride,
                 int64_t tabLen, V* z, int64_t zStride, int batch, hipStream_t s);
template <typename V>
void launch_rhs_sines(const SumItem* items, int nitems, int d, int nterms, const double* eta, const int32_t* m, V* b,
                      int64_t bStride, int batch, hipStream_t s);
template <typename V>
void launch_cg_init(const V* f, V* r, V* h, V* dv, const V* D, int64_t n, int batch, double* partials, int nb,
                    hipStream_t s);
// epsInst: per-instance tolerances (device array) or null for eps
void launch_cg_start(CGState* st, const double* partials, int nb, double eps, const double* epsInst, int rule,
                     int maxit, int batch, hipStream_t s);
template <typename V>
void launch_cg_iter_tail(V* r, void* x, bool x64, V* h, V* dv, const V* g, const V* D, int64_t n, int batch,
                         double* partials, int nb, CGState* st, hipStream_t s);
void launch_residual(const double* f, const double* Ax, double* r, int64_t n, int batch, double* partials, int nb,
                     hipStream_t s);
void launch_cast_d2f(const double* a, float* b, int64_t total, hipStream_t s);
void launch_cast_f2d(const float* a, double* b, int64_t total, hipStream_t s);
void launch_axpy_correction(double* x, const float* e, int64_t total, hipStream_t s);
void launch_zero(void* p, int64_t bytes, hipStream_t s);
int cg_partial_blocks(int64_t n);

}  // namespace lsg
