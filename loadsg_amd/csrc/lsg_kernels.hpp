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
    const BoxShape* shapes;   // box shapes of the apply items (ApplyItem::shp)
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
    int batch;            // instance slots in the grid (workgroup = family-major, slot fastest)
    const CGState* st;    // may be null
    const int* map;       // slot -> instance (null: slot = instance)
};

// Device copies of the sparse-path gather tables (SparseTables, lsg_plan.hpp).
struct SparseDev {
    int d, nd;
    const int64_t* nodPtr;
    const int32_t* nodCol;
    const double* nodVal;
    const int64_t* prolOff;
    const int32_t* prolAB;
    const int64_t* cellOff;
    const int32_t* cellCorner;
    const int32_t* cellLo;
    const int64_t* cellColOff;
    const int32_t* cellCol;
    const int64_t* rcPtr;
    const int32_t* rcCell;
    const int32_t* rcCorner;
    const int32_t* rcY;
    const int64_t* restOff;
    const int32_t* restSLR;
    const int64_t* convPtr;
    const int32_t* convCol;
    const double* convVal;
    // convVal / nodVal rounded to FP32 for the FP32 kernels (half the table traffic; float(w)
    // is what those kernels formed from the FP64 table anyway, so results do not change)
    const float* convValF;
    const float* nodValF;
    // 16-bit copies of the cell-wise apply's index tables (null unless every box and every
    // depth's cell list has fewer than 32768 entries)
    const int16_t* cellCorner16;
    const int16_t* cellCol16;
    const int16_t* restSLR16;
    const int16_t* convCol16;
    const int32_t* pos;
    const int32_t* tOfDev;
    const int32_t* depthT;
    const int64_t* boxOff;
    const int64_t* boxSize;
    const int64_t* blkOff;
    const int64_t* blkSize;
};

template <typename V>
struct SparseApplyArgs {
    SparseDev sp;
    const FamilyItem* fam;    // single-item families of one wavefront level
    const ApplyItem* items;
    int nA, nC;
    const V* coefTab;
    int64_t coefStride, tabLen;
    const V* nodal;
    int64_t nStride;
    const V* cbuf;
    int64_t cStride;
    V* wbuf;
    int64_t wStride;
    V* part;
    int64_t pStride;
    V* zbuf;
    int64_t zStride;
    int batch;
    const CGState* st;
    const int* map;           // slot -> instance (null: slot = instance)
    const V* cellCoef;        // per instance: a at the centre of every cell of depth T+1, then c
    int64_t ccStride;
    // stored element matrices (null: element matrices from cellCoef): entry e of cell c of
    // instance i at cellMat[i * matStride + e * nCells + c], e = cellEntry(p, q)
    const V* cellMat;
    int64_t matStride, nCells;
    // cell-wise kernels (element matrices from the cell coefficients, d <= SP_CELL_MAXD):
    // values per instance of the LDS scratch of this launch (0: scratch in zbuf)
    int64_t ldsStride;
};

// Position of the pair of cell corners p <= q (bit i = upper end in dimension i) among the
// 2^(d-1) (2^d + 1) entries of a symmetric element matrix, in the order of the reference's
// nested CellIndexIterator loops (cells/iterator_neu.cpp:34-49, LocalStiffnessMatrices.cpp:463-500).
__host__ __device__ inline int cellEntry(unsigned p, unsigned q, unsigned nc) {
    return (int)(p * nc - (p * (p - 1)) / 2 + (q - p));
}

template <typename V>
void launch_sp_cellcoef(const SparseDev& s, int64_t nCells, int nA, int nC, const V* coefTab, int64_t coefStride,
                        int64_t tabLen, V* cc, int64_t ccStride, int batch, hipStream_t str);
template <typename V>
void launch_sp_nodal(const SparseDev& s, const V* x, int64_t xStride, V* nodal, int64_t nStride, int64_t nml,
                     int batch, const CGState* st, hipStream_t str, const int* map = nullptr);
template <typename V>
void launch_sp_prol(const ProlItem* items, int nitems, const SparseDev& s, const V* nodal, int64_t nStride, V* cbuf,
                    int64_t cStride, int batch, const CGState* st, hipStream_t str, const int* map = nullptr);
template <typename V>
void launch_sp_apply(const SparseApplyArgs<V>& a, int nitems, hipStream_t str, int cls);
// scratch classes of the cell-wise sparse apply (one launch per class and wavefront level)
enum SpClass : int { SP_SMALL = 0, SP_MEDIUM = 1, SP_LARGE = 2, SP_GLOBAL = 3 };
constexpr int SP_NCLASS = 4;
template <typename V>
void launch_sp_diag(const SparseDev& s, const uint8_t* activeDev, int n, const V* cellCoef, int64_t ccStride,
                    const V* cellMat, int64_t matStride, int64_t nCells, V* z, int64_t zStride, int batch,
                    hipStream_t str);

template <typename V>
void launch_nodal(const NodalItem* items, int nitems, int d, const V* x, int64_t xStride, V* nodal, int64_t nStride,
                  int64_t maxBox, int batch, const CGState* st, hipStream_t s, const int* map = nullptr);
template <typename V>
void launch_prol(const ProlItem* items, int nitems, int d, const V* nodal, int64_t nStride, V* cbuf, int64_t cStride,
                 int64_t maxBox, int64_t maxChain, int batch, const CGState* st, hipStream_t s,
                 const int* map = nullptr);
template <typename V>
int launch_apply(const ApplyArgs<V>& a, int nitems, int batch, int threads, size_t lds, hipStream_t s);
template <typename V>
void launch_sum(const SumItem* items, int nitems, const int64_t* list, int d, const V* part, int64_t pStride, V* y,
                int64_t yStride, int batch, const CGState* st, hipStream_t s, const int* map = nullptr);
template <typename V>
void launch_diag(const SumItem* items, int nitems, int d, int nA, int nC, const V* coefTab, int64_t coefStride,
                 int64_t tabLen, V* z, int64_t zStride, int batch, hipStream_t s);
template <typename V>
void launch_rhs_tab(const SumItem* items, int nitems, int d, int nterms, const double* w, const double* tab,
                    int64_t tabLen, V* b, int64_t bStride, int batch, hipStream_t s, const int32_t* pos,
                    const uint8_t* active);
template <typename V>
void launch_rhs_sines(const SumItem* items, int nitems, int d, int nterms, const double* eta, const int32_t* m, V* b,
                      int64_t bStride, int batch, hipStream_t s, const int32_t* pos = nullptr,
                      const uint8_t* active = nullptr);
// Depth-block preconditioner of a CG solve (BlockPre, lsg_plan.hpp) on the device: h = s .*
// (B_T^-1 (s .* r)) on every depth block T with an inverse, s = sqrt(d0 / D), h = r / D on the
// others; part receives the r.h partial of every (instance, block).  items == null: the
// diagonal preconditioner.
constexpr int PRE_MM_MIN = 32;
template <typename V>
struct BlockPreArgs {
    const PreBlock* items = nullptr;
    int nblk = 0;
    int maxN = 0;           // largest block with an inverse (LDS of the kernel)
    const V* inv = nullptr;
    const V* d0 = nullptr;
    double* part = nullptr;
    // fast diagonalisation of the larger blocks (BlockPre::fdV / fdLam / fdList / fdPerm)
    int d = 0;
    int nfd = 0, maxFd = 0; // blocks on that path, the largest product set of them
    const int32_t* fdList = nullptr;
    const int32_t* fdPerm = nullptr;
    // FP32: dense blocks of at least PRE_MM_MIN points on the matrix cores (blockmm_kernel)
    int nmm = 0, mmMax = 0, mmGroups = 1;
    const int32_t* mmList = nullptr;
    const V* fdV = nullptr;
    const V* fdLam = nullptr;
};
template <typename V>
void launch_cg_init(const V* f, V* r, V* h, V* dv, const V* D, int64_t n, int batch, double* partials, int nb,
                    hipStream_t s, const BlockPreArgs<V>& pre = BlockPreArgs<V>());
// epsInst: per-instance tolerances (device array) or null for eps; bpart / nbb: r.h partials
// of the depth-block preconditioner (null: slot 0 of partials)
void launch_cg_start(CGState* st, const double* partials, int nb, double eps, const double* epsInst, int rule,
                     int maxit, int batch, hipStream_t s, const double* bpart = nullptr, int nbb = 0);
template <typename V>
void launch_cg_iter_tail(V* r, void* x, bool x64, V* h, V* dv, const V* g, const V* D, int64_t n, int batch,
                         double* partials, int nb, CGState* st, hipStream_t s,
                         const BlockPreArgs<V>& pre = BlockPreArgs<V>());
// r = f - Ax and the r.r, f.f partials for `batch` instances (slots map[0..batch) if map)
void launch_residual(const double* f, const double* Ax, double* r, int64_t n, int batch, double* partials, int nb,
                     hipStream_t s, const int* map = nullptr);
void launch_cast_d2f(const double* a, float* b, int64_t total, hipStream_t s);
void launch_cast_f2d(const float* a, double* b, int64_t total, hipStream_t s);
void launch_axpy_correction(double* x, const float* e, int64_t total, hipStream_t s);
void launch_zero(void* p, int64_t bytes, hipStream_t s);
void launch_fill_one(void* p, int64_t n, bool fp64, hipStream_t s);   // p[i] = 1 (double or float)
// rows rows[0..nrows) (device list) of a row-major (., n) FP64 array set to zero
void launch_zero_rows(double* a, int64_t n, const int* rows, int nrows, hipStream_t s);
int cg_partial_blocks(int64_t n);
// basis transformations (lsg_transform.hip); mode 2 prewavelets of one |T|_1 level,
// 3 hierarchical representation of prewavelets
void launch_tr_box(int mode, const TrItem* items, const int32_t* list, int nitems, const int32_t* gidx,
                   const double* fac, int64_t facTot, int64_t maxBox, const double* in, double* out, double* scratch,
                   int64_t n, int64_t nml, int batch, hipStream_t s);
void launch_tr_gather(const int32_t* rowDev, const int64_t* rowPtr, const int32_t* col, int64_t nrows,
                      const double* scratch, double* v, int64_t n, int64_t nml, int sub, int batch, hipStream_t s);
// adaptive grids: mode 0 one depth of calcPrewByNodal (nitems 1), 1 hierarchical
// representation of the prewavelets of every depth (nitems = number of depths)
void launch_tr_adapt(int mode, const TrItem* items, int depth, int nitems, int64_t maxBox, const TrComp* comps,
                     const int32_t* compOff, const int32_t* gidx, const uint8_t* bstat, const double* afac,
                     int64_t afacTot, const double* src, double* hier, double* out, double* scratch, int64_t n,
                     int64_t boxTot, int first, int64_t ldsLimit, int batch, hipStream_t s);
void launch_tr_pole(const int32_t* poles, const double* in, double* out, int64_t n, int batch, hipStream_t s);
void launch_tr_dehier(const int32_t* list, int64_t cnt, const int32_t* poles, double* v, int64_t n, int batch,
                      hipStream_t s);
void launch_tr_energy(const double* x, const double* y, int64_t n, double* out, int batch, hipStream_t s);
// per instance: out[2b] = max |x| over all points, out[2b+1] = sqrt(sum over active x^2 / dofs)
template <typename V>
void launch_vec_norms(const V* x, int64_t n, const uint8_t* active, double dofs, double* out, int batch,
                      hipStream_t s);

}  // namespace lsg
