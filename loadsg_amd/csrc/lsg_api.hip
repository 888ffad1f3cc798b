// lsg_api.hip -- C ABI (include/loadsg_amd.h): device memory, the stage sequence
// of one matvec, and the CG / iterative-refinement drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <chrono>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/loadsg_amd.h"
#include "lsg_kernels.hpp"

using namespace lsg;

namespace {
thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
}  // namespace

namespace lsg {
// error reporting shared with the other translation units of the C ABI
int api_fail(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace lsg

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(LSG_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct lsg_grid {
    Grid g;
    // the same grid on the cell-wise (sparse) path, built on first use by an operator with
    // a general coefficient (LSG_COEF_CELL / LSG_COEF_FUNC) or by lsg_grid_cells; it keeps
    // g's device order
    std::unique_ptr<Grid> cellG;
};

namespace {

struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = std::max<size_t>(b, 16);
        return hipMalloc(&p, bytes);
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Pinned host memory (hipHostMalloc): targets of asynchronous copies that must not
// serialise the host thread.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    hipError_t alloc(size_t b) {
        if (p) { (void)hipHostFree(p); p = nullptr; }
        bytes = std::max<size_t>(b, 16);
        return hipHostMalloc(&p, bytes, hipHostMallocDefault);
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Separable representation of the coefficient families: a = sum_m prod_j g_mj(x_j),
// c = sum_m prod_j h_mj(x_j), sampled at the 1D cell centres (2c+1)/2^(t+1) of every
// level t.  Table of one (term, dim): level t occupies [2^t - 2, 2^(t+1) - 2).
bool matCoefKind(int kind) { return kind == LSG_COEF_MATRIX || kind == LSG_COEF_INTEGRATOR; }
bool cellCoefKind(int kind) { return kind == LSG_COEF_CELL || kind == LSG_COEF_FUNC || matCoefKind(kind); }
int64_t cellEntries(int d) { return (int64_t(1) << (d - 1)) * ((int64_t(1) << d) + 1); }

// Dense grids run on their single-item plan (every prolongation dimension materialised)
// for batches up to singlePlanMaxBatch(): measured faster than the families along the last
// prolongation dimension at every batch from 1 to 512 at d=5 level 7 (FP64: 0.21 vs
// 0.52 ms at batch 1, 31.3 vs 32.8 ms at 512).  LSG_SINGLE_PLAN_BATCH=0 selects the family
// plan.  Batches up to oneClassMaxBatch() run one apply launch per wavefront level
// (LSG_ONE_CLASS_BATCH overrides).
int envInt(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}
int singlePlanMaxBatch() { return envInt("LSG_SINGLE_PLAN_BATCH", 1 << 30); }
int oneClassMaxBatch() { return envInt("LSG_ONE_CLASS_BATCH", 4); }
// LDS bytes the adaptive-grid transform may give one workgroup's boxes (default: the
// device's limit); larger boxes are staged in global memory
int64_t trLdsLimit(int64_t deviceMax) { return std::min<int64_t>(deviceMax, envInt("LSG_TR_LDS_BYTES", (int)deviceMax)); }

int coefTerms(const lsg_coef& c, int d, int& nA, int& nC) {
    switch (c.kind) {
        case LSG_COEF_CELL: nA = c.cell_a != nullptr; nC = c.cell_c != nullptr; break;
        case LSG_COEF_FUNC: nA = c.a_fn != nullptr; nC = c.c_fn != nullptr; break;
        case LSG_COEF_MATRIX: nA = c.cell_mat != nullptr; nC = 0; break;
        case LSG_COEF_INTEGRATOR: nA = c.integ != nullptr; nC = 0; break;
        case LSG_COEF_CONST: nA = c.a0 != 0.0; nC = c.c0 != 0.0; break;
        case LSG_COEF_CFG2: nA = 2; nC = 2; break;
        case LSG_COEF_CFG3: nA = 1 + d; nC = 0; break;
        case LSG_COEF_EXPSIN: nA = 1; nC = 0; if (c.nxi < 0 || (c.nxi > 0 && !c.xi)) return -1; break;
        default: return -1;
    }
    return 0;
}

void coefTables(const lsg_coef& c, int d, int maxLevel, int64_t tabLen, std::vector<double>& tab) {
    int nA, nC;
    coefTerms(c, d, nA, nC);
    const int nt = nA + nC;
    tab.assign((size_t)nt * d * tabLen, 1.0);
    if (cellCoefKind(c.kind)) return;   // sampled per cell instead (cellValues)
    auto fill = [&](int term, int dim, auto fn) {
        double* t0 = &tab[((size_t)term * d + dim) * tabLen];
        for (int t = 1; t <= maxLevel; ++t)
            for (int cc = 0; cc < (1 << t); ++cc) {
                double x = (2.0 * cc + 1.0) / double(1 << (t + 1));
                t0[(1 << t) - 2 + cc] = fn(x);
            }
    };
    switch (c.kind) {
        case LSG_COEF_CONST: {
            int term = 0;
            if (c.a0 != 0.0) { fill(term, 0, [&](double) { return c.a0; }); ++term; }
            if (c.c0 != 0.0) { fill(term, 0, [&](double) { return c.c0; }); ++term; }
            break;
        }
        case LSG_COEF_CFG2:   // a = 1 + 0.5 prod cos(pi x_i); c = 1 + x_1
            for (int i = 0; i < d; ++i)
                fill(1, i, [&](double x) { return (i == 0 ? 0.5 : 1.0) * std::cos(M_PI * x); });
            fill(3, 0, [&](double x) { return x; });
            break;
        case LSG_COEF_CFG3:   // a = 1 + sum x_i^2
            for (int i = 0; i < d; ++i) fill(1 + i, i, [&](double x) { return x * x; });
            break;
        case LSG_COEF_EXPSIN:  // a = prod_j exp(sum_{k mod d = j} xi_k sin(k pi x_j) / k)
            for (int j = 0; j < d; ++j)
                fill(0, j, [&](double x) {
                    double s = 0.0;
                    for (int k = 1; k <= c.nxi; ++k)
                        if (k % d == j) s += c.xi[k - 1] * std::sin(double(k) * M_PI * x) / double(k);
                    return std::exp(s);
                });
            break;
    }
}

// Grid of the cell-wise path: the grid itself when it is sparse, else its cellG copy.
int cellGrid(lsg_grid* grid, const Grid*& out) {
    if (grid->g.sparse) {
        out = &grid->g;
        return LSG_OK;
    }
    if (!grid->cellG) {
        std::unique_ptr<Grid> cg(new Grid());
        std::string err;
        const Grid& g = grid->g;
        if (int rc = buildGrid(g.d, g.n, g.idx.data(), g.active.data(), *cg, err, true)) return fail(rc, err);
        grid->cellG = std::move(cg);
    }
    out = grid->cellG.get();
    return LSG_OK;
}

// Centre of every cell of depth T+1 of every depth (cell k of depth t: lower position
// cellLo, level T_i: x_i = (lo_i + 1/2) / 2^T_i), cells in table order.
void cellCentres(const Grid& g, std::vector<double>& x, std::vector<int32_t>* depthOf) {
    const SparseTables& S = g.sp;
    const int d = g.d;
    const int nd = (int)g.depths.size();
    const int64_t nCells = S.cellOff[nd];
    x.resize((size_t)nCells * d);
    if (depthOf) depthOf->resize((size_t)nCells);
    for (int t = 0; t < nd; ++t)
        for (int64_t c = S.cellOff[t]; c < S.cellOff[t + 1]; ++c) {
            for (int i = 0; i < d; ++i)
                x[(size_t)c * d + i] = (double(S.cellLo[(size_t)c * d + i]) + 0.5) / double(1 << g.depths[t][i]);
            if (depthOf) (*depthOf)[(size_t)c] = t;
        }
}

// (a, c) per cell, interleaved, for a general coefficient
int cellValues(const Grid& g, const lsg_coef& c, std::vector<double>& ac) {
    const int d = g.d;
    std::vector<double> x;
    cellCentres(g, x, nullptr);
    const int64_t nCells = (int64_t)(x.size() / d);
    ac.assign((size_t)2 * nCells, 0.0);
    for (int64_t k = 0; k < nCells; ++k) {
        double a = 0.0, cc = 0.0;
        if (c.kind == LSG_COEF_CELL) {
            if (c.cell_a) a = c.cell_a[k];
            if (c.cell_c) cc = c.cell_c[k];
        } else {
            if (c.a_fn) a = c.a_fn(&x[(size_t)k * d], c.ctx);
            if (c.c_fn) cc = c.c_fn(&x[(size_t)k * d], c.ctx);
        }
        if (!std::isfinite(a) || !std::isfinite(cc)) return fail(LSG_ERR_NAN, "non-finite coefficient value");
        ac[(size_t)2 * k] = a;
        ac[(size_t)2 * k + 1] = cc;
    }
    return LSG_OK;
}

// Stored element matrices of every cell for LSG_COEF_MATRIX / LSG_COEF_INTEGRATOR, laid out
// entry-major (entry e of cell k at e * nCells + k: the cells a wave handles are adjacent in
// memory for every entry).  Pairs with a boundary corner are left 0 and never read.
int cellMatrices(const Grid& g, const lsg_coef& c, std::vector<double>& E) {
    const int d = g.d;
    const SparseTables& S = g.sp;
    const int nd = (int)g.depths.size();
    const int64_t nCells = S.cellOff[nd];
    const unsigned nc = 1u << d;
    const int64_t ne = cellEntries(d);
    E.assign((size_t)(nCells * ne), 0.0);
    int32_t level[MAXD], pos[MAXD];
    for (int t = 0; t < nd; ++t) {
        for (int i = 0; i < d; ++i) level[i] = g.depths[t][i];
        for (int64_t k = S.cellOff[t]; k < S.cellOff[t + 1]; ++k) {
            const int32_t* corner = &S.cellCorner[(size_t)(k * nc)];
            for (int i = 0; i < d; ++i) pos[i] = S.cellLo[(size_t)(k * d + i)];
            for (unsigned p = 0; p < nc; ++p) {
                if (corner[p] < 0) continue;
                for (unsigned q = p; q < nc; ++q) {
                    if (corner[q] < 0) continue;
                    const int e = cellEntry(p, q, nc);
                    const double v = c.kind == LSG_COEF_MATRIX ? c.cell_mat[(size_t)(k * ne + e)]
                                                               : c.integ(d, level, pos, p, q, c.ctx);
                    if (!std::isfinite(v)) return fail(LSG_ERR_NAN, "non-finite element matrix entry");
                    E[(size_t)(e * nCells + k)] = v;
                }
            }
        }
    }
    return LSG_OK;
}

// ---- separable load vectors: 1D moments on the host
// 4-point Gauss-Legendre rule on [-1, 1] (abscissae and weights of the Legendre
// polynomial P_4), applied on every cell of the finest 1D level of the grid: every
// level-t hat and prewavelet is linear there.
constexpr double GL_X[4] = {-0.861136311594052575223946488893, -0.339981043584856264802665759103,
                            0.339981043584856264802665759103, 0.861136311594052575223946488893};
constexpr double GL_W[4] = {0.347854845137453857373063949222, 0.652145154862546142626936050778,
                            0.652145154862546142626936050778, 0.347854845137453857373063949222};

void rhsNodes(int maxLevel, std::vector<double>& x) {
    const int64_t cells = int64_t(1) << maxLevel;
    x.resize((size_t)cells * 4);
    for (int64_t c = 0; c < cells; ++c)
        for (int q = 0; q < 4; ++q) x[(size_t)c * 4 + q] = (double(c) + 0.5 * (1.0 + GL_X[q])) / double(cells);
}

// moments m(t, k) = int g psi_{t,k} of one factor given at the rhsNodes, for every level
// t <= maxLevel and odd k: first the level-t hat moments, then the prewavelet stencil;
// entry 2^(t-1) - 1 + (k - 1) / 2 of tab (length 2^maxLevel - 1)
void factorMoments(int maxLevel, const double* g, double* tab) {
    const int64_t cells = int64_t(1) << maxLevel;
    const double hf = 1.0 / double(cells);
    std::vector<double> H;
    for (int t = 1; t <= maxLevel; ++t) {
        const int64_t nJ = (int64_t(1) << t) - 1, span = int64_t(1) << (maxLevel - t);
        H.assign((size_t)nJ + 2, 0.0);
        for (int64_t J = 1; J <= nJ; ++J) {
            double acc = 0.0;
            for (int64_t c = (J - 1) * span; c < (J + 1) * span; ++c)
                for (int q = 0; q < 4; ++q) {
                    const double x = (double(c) + 0.5 * (1.0 + GL_X[q])) * hf;
                    const double hat = 1.0 - std::fabs(x * double(int64_t(1) << t) - double(J));
                    acc += 0.5 * hf * GL_W[q] * g[(size_t)c * 4 + q] * hat;
                }
            H[(size_t)J] = acc;
        }
        for (int64_t k = 1; k <= nJ; k += 2) {
            int32_t pos[5];
            double co[5];
            const int m = prewaveletStencil1D((int32_t)k, t, pos, co);
            double v = 0.0;
            for (int e = 0; e < m; ++e) v += co[e] * H[(size_t)pos[e]];
            tab[(int64_t(1) << (t - 1)) - 1 + (k - 1) / 2] = v;
        }
    }
}

}  // namespace

namespace {
// Device copies of the basis-transformation tables and their work vectors.
struct TrDev {
    TrTables host;
    DevMem items, gidx, list, rowDev, rowPtr, col, fac, poles, byLevel;
    DevMem bstat, comps, compOff, afac;   // adaptive grids
    DevMem hier, work, work2, scratch, norms;
};
}  // namespace

struct lsg_op {
    lsg_grid* grid = nullptr;
    const Grid* G = nullptr;   // the grid the kernels run on: grid->g, or its cell-wise copy
    const Plan* P = nullptr;   // its plan: G->plan, or the single-item plan of a dense grid
    bool cellKind = false;     // general coefficient given per cell (cellCoef filled on the host)
    bool matKind = false;      // stored element matrices per cell (cellMat filled on the host)
    int device = 0;            // HIP device the operator's memory and stream belong to
    int prec = 64;
    size_t vs = 8;
    int batch = 1;
    int kind = 0;
    int nA = 0, nC = 0;
    int64_t tabLen = 0;
    bool general = false;
    int64_t maxW = 0;
    // apply launches: per wavefront level, families bucketed by their largest box
    struct ApplyLaunch {
        int level, cls;
        int count;
        int64_t maxBox;     // largest box of the bucket (LDS box size)
        int64_t maxW;       // largest weight table of the bucket
        int G;              // batch instances per workgroup
        std::unique_ptr<DevMem> fam;
    };
    // three bucketings of the apply families by box size (12 fine classes, 5 coarse, one):
    // the autotuner times each with its best workgroup sizes and keeps the fastest
    struct ClassSet {
        std::vector<ApplyLaunch> launches;
        int threadsOf[16] = {};   // per size class; 0 = autotune on first use
        int nClasses = 0;
    };
    ClassSet sets[3];
    int cs = 0;                 // active set
    ClassSet& act() { return sets[cs]; }
    const ClassSet& act() const { return sets[cs]; }
    int64_t maxLds = 0;
    DevMem dNodalItems, dSumItems, dSumList, dApplyItems, dShapes;
    std::vector<std::unique_ptr<DevMem>> dProl;   // per sweep dim
    std::vector<int64_t> prolMax;                 // largest box per sweep dim
    std::vector<int64_t> prolChain;               // most values in one chain per sweep dim
    DevMem coefTab, nodal, cbuf, wbuf, part, xs, ys;
    // sparse path: gather tables, per-level apply families, scratch
    bool sparse = false;
    std::vector<std::unique_ptr<DevMem>> spTables;
    SparseDev sd{};
    // sparse apply launches: per wavefront level and scratch class (SpClass) the families,
    // their count and the LDS values per instance
    struct SpLaunch {
        int count;
        int cls;
        int64_t ldsStride;
        std::unique_ptr<DevMem> fam;
    };
    std::vector<SpLaunch> spLevels;
    DevMem zbuf, activeDev, cellCoef;
    DevMem cellMat;            // stored element matrices: [instance][entry][cell]
    int64_t nEnt = 0;          // entries per cell
    // instances still iterating in a CG solve: the matvec grids cover only these slots
    DevMem actMap;
    int nMap = 0;             // > 0: launch nMap slots mapped through actMap
    int64_t nCells = 0;
    bool cellCoefValid = false;
    // solver buffers
    DevMem r, h, dv, g, D, f, x, tmp, partials, st, aux;
    DevMem epsInst, rows;        // iterative refinement: per-instance inner tolerances, rows to clear
    DevMem norms2;               // lsg_vector_norms results (2 per instance)
    DevMem ones;                 // D = 1 of the unpreconditioned CG
    PinnedBuf hostSt, hostMap, hostPart;   // host copies of the CG state, instance lists, residual partials
    DevMem rhsEta, rhsM;         // lsg_rhs_sines parameters (grown on demand)
    bool solverReady = false, diagValid = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    bool timed = false;
    // CG iteration chunk captured as a hipGraph (while every instance is launched):
    // valid for the x pointer / precision / chunk length it was captured with
    hipGraphExec_t chunkExec = nullptr;
    const void* chunkX = nullptr;
    const void* chunkD = nullptr;   // diagonal the chunk divides by (D, or ones when unpreconditioned)
    int chunkK = 0;
    bool chunkX64 = false;
    bool chunkPre = false;      // chunk captured with the depth-block preconditioner
    // preconditioner of lsg_cg_solve and of the inner solves of lsg_ir_solve
    // (LSG_PRECOND_DIAG: the reference's precond3; LSG_PRECOND_DEPTH_BLOCK: BlockPre)
    int precKind = LSG_PRECOND_DIAG;
    DevMem preItems, preInv, preD0, prePart, preFdList, preFdPerm, preFdV, preFdLam, preMmList;
    int preNblk = 0, preMaxN = 0, preNfd = 0, preMaxFd = 0, preNmm = 0, preMmMax = 0;
    bool capturing = false;   // matvec launches record no timing events while captured
    std::unique_ptr<TrDev> tr;   // basis transformations (built on first use)
    int64_t n() const { return G->n; }
};

namespace {

// Sparse path: a, then c, at the centre of every cell, per instance, from the coefficient
// tables; the stencil and the diagonal read these.  Operators with a general coefficient
// receive them from the host (lsg_op_set_coef) and never run this.
int ensureCellCoef(lsg_op* op) {
    if (!op->sparse || op->cellCoefValid) return LSG_OK;
    if (op->capturing) return fail(LSG_ERR_ARG, "cell coefficients must be built before graph capture");
    const int64_t cs = (int64_t)(op->nA + op->nC) * op->G->d * op->tabLen;
    if (op->prec == 64)
        launch_sp_cellcoef<double>(op->sd, op->nCells, op->nA, op->nC, op->coefTab.as<double>(), cs, op->tabLen,
                                   op->cellCoef.as<double>(), 2 * op->nCells, op->batch, op->stream);
    else
        launch_sp_cellcoef<float>(op->sd, op->nCells, op->nA, op->nC, op->coefTab.as<float>(), cs, op->tabLen,
                                  op->cellCoef.as<float>(), 2 * op->nCells, op->batch, op->stream);
    HIPCHK(hipGetLastError());
    op->cellCoefValid = true;
    return LSG_OK;
}

// nSub > 0: launch only the first nSub slots listed in actMap (a residual over the
// instances still unconverged); otherwise, inside a CG solve, the compacted CG list.
template <typename V>
int runMatvecSparse(lsg_op* op, const V* x, V* y, const CGState* st, int nSub) {
    const Grid& g = *op->G;
    const Plan& P = *op->P;
    const int d = g.d;
    const int nm = nSub > 0 ? nSub : (st ? op->nMap : 0);
    const int* map = nm > 0 ? op->actMap.as<int>() : nullptr;
    const int B = map ? nm : op->batch;
    hipStream_t s = op->stream;
    if (int rc = ensureCellCoef(op)) return rc;   // (done before a CG chunk is captured)
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[0], s));
    launch_sp_nodal<V>(op->sd, x, g.n, op->nodal.as<V>(), g.nml, g.nml, B, st, s, map);
    for (int dd = 0; dd < d; ++dd)
        launch_sp_prol<V>(op->dProl[dd]->as<ProlItem>(), (int)P.prol[dd].size(), op->sd, op->nodal.as<V>(), g.nml,
                          op->cbuf.as<V>(), P.caseSize, B, st, s, map);
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[1], s));
    SparseApplyArgs<V> a{};
    a.sp = op->sd;
    a.items = op->dApplyItems.as<ApplyItem>();
    a.nA = op->nA;
    a.nC = op->nC;
    a.coefTab = op->coefTab.as<V>();
    a.coefStride = (int64_t)(op->nA + op->nC) * d * op->tabLen;
    a.tabLen = op->tabLen;
    a.nodal = op->nodal.as<V>();
    a.nStride = g.nml;
    a.cbuf = op->cbuf.as<V>();
    a.cStride = P.caseSize;
    a.wbuf = op->wbuf.as<V>();
    a.wStride = P.wSize;
    a.part = op->part.as<V>();
    a.pStride = P.partSize;
    a.zbuf = op->zbuf.as<V>();
    a.zStride = P.zSize;
    a.batch = B;
    a.st = st;
    a.map = map;
    a.cellCoef = op->cellCoef.as<V>();
    a.ccStride = 2 * op->nCells;
    a.cellMat = op->matKind ? op->cellMat.as<V>() : nullptr;
    a.matStride = op->nCells * op->nEnt;
    a.nCells = op->nCells;
    for (auto& L : op->spLevels) {
        a.fam = L.fam->as<FamilyItem>();
        a.ldsStride = L.ldsStride;
        launch_sp_apply<V>(a, L.count, s, L.cls);
    }
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[2], s));
    launch_sum<V>(op->dSumItems.as<SumItem>(), (int)P.sum.size(), op->dSumList.as<int64_t>(), d, op->part.as<V>(),
                  P.partSize, y, g.n, B, st, s, map);
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[3], s));
    HIPCHK(hipGetLastError());
    op->timed = true;
    return LSG_OK;
}

template <typename V>
int runMatvec(lsg_op* op, const V* x, V* y, const CGState* st, int nSub = 0) {
    if (op->sparse) return runMatvecSparse<V>(op, x, y, st, nSub);
    const Grid& g = *op->G;
    const Plan& P = *op->P;
    const int d = g.d;
    // inside a CG solve with a compacted instance list, or for a residual over a
    // subset (nSub > 0), launch only the slots listed in actMap
    const int nm = nSub > 0 ? nSub : (st ? op->nMap : 0);
    const int* map = nm > 0 ? op->actMap.as<int>() : nullptr;
    const int B = map ? nm : op->batch;
    hipStream_t s = op->stream;
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[0], s));
    launch_nodal<V>(op->dNodalItems.as<NodalItem>(), (int)P.nodal.size(), d, x, g.n, op->nodal.as<V>(), g.nml,
                    P.maxBox, B, st, s, map);
    for (int dd = 0; dd < d; ++dd)
        launch_prol<V>(op->dProl[dd]->as<ProlItem>(), (int)P.prol[dd].size(), d, op->nodal.as<V>(), g.nml,
                       op->cbuf.as<V>(), P.caseSize, op->prolMax[dd], op->prolChain[dd], B, st, s, map);
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[1], s));
    ApplyArgs<V> a{};
    a.d = d;
    a.general = op->general ? 1 : 0;
    a.nA = op->nA;
    a.nC = op->nC;
    a.coefTab = op->coefTab.as<V>();
    a.coefStride = (int64_t)(op->nA + op->nC) * d * op->tabLen;
    a.tabLen = op->tabLen;
    a.nodal = op->nodal.as<V>();
    a.nStride = g.nml;
    a.cbuf = op->cbuf.as<V>();
    a.cStride = P.caseSize;
    a.wbuf = op->wbuf.as<V>();
    a.wStride = P.wSize;
    a.part = op->part.as<V>();
    a.pStride = P.partSize;
    a.st = st;
    a.items = op->dApplyItems.as<ApplyItem>();
    a.shapes = op->dShapes.as<BoxShape>();
    a.batch = B;
    a.map = map;
    for (auto& L : op->act().launches) {
        a.fam = L.fam->as<FamilyItem>();
        a.maxBox = L.maxBox;
        a.maxW = L.maxW;
        a.G = L.G;
        const int thr = op->act().threadsOf[L.cls] ? op->act().threadsOf[L.cls] : 256;
        const size_t lds = ((op->general ? 4 : 2) * L.maxBox + L.maxW) * sizeof(V);
        if (launch_apply<V>(a, L.count, B, thr, lds, s) != 0)
            return fail(LSG_ERR_UNSUPPORTED, "depth box too large for the chosen block size");
    }
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[2], s));
    launch_sum<V>(op->dSumItems.as<SumItem>(), (int)P.sum.size(), op->dSumList.as<int64_t>(), d, op->part.as<V>(),
                  P.partSize, y, g.n, B, st, s, map);
    if (!op->capturing) HIPCHK(hipEventRecord(op->ev[3], s));
    HIPCHK(hipGetLastError());
    op->timed = true;
    return LSG_OK;
}

int matvecDev(lsg_op* op, const void* x, void* y, const CGState* st) {
    if (op->prec == 64) return runMatvec<double>(op, (const double*)x, (double*)y, st);
    return runMatvec<float>(op, (const float*)x, (float*)y, st);
}

int64_t classMax(const lsg_op* op, int cls) {
    int64_t m = 0;
    for (auto& L : op->act().launches)
        if (L.cls == cls) m = std::max(m, L.maxBox);
    return m;
}

// LDS bytes of one apply workgroup (one family of one instance)
int64_t applyLdsBytes(const lsg_op* op, int64_t maxBox, int64_t maxW, int threads) {
    (void)threads;
    // general: U, Y, Z, Zt; separable: two boxes (P parked between them)
    return ((op->general ? 4 : 2) * maxBox + maxW) * (int64_t)op->vs;
}

bool threadsFit(const lsg_op* op, int cls, int threads) {
    for (auto& L : op->act().launches)
        if (L.cls == cls && (applyLdsBytes(op, L.maxBox, L.maxW, threads) > op->maxLds ||
                             (L.maxBox + threads - 1) / threads > 16))
            return false;
    return true;
}

// apply-stage time (ms) of the active class set with its current workgroup sizes
int timeApply(lsg_op* op, float& ms) {
    float tot = 0.f;
    for (int rep = 0; rep < 3; ++rep) {
        int rc = matvecDev(op, op->xs.p, op->ys.p, nullptr);
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(op->ev[3]));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, op->ev[1], op->ev[2]));
        if (rep > 0) tot += t;
    }
    ms = tot;
    return LSG_OK;
}

int chooseThreads(lsg_op* op, int threads) {
    const int cands[8] = {64, 128, 192, 256, 320, 384, 448, 512};
    if (op->chunkExec) {          // a captured CG chunk holds the old launch shapes
        HIPCHK(hipGraphExecDestroy(op->chunkExec));
        op->chunkExec = nullptr;
    }
    if (threads) {
        for (int c = 0; c < op->act().nClasses; ++c) {
            if (!threadsFit(op, c, threads)) return fail(LSG_ERR_UNSUPPORTED, "apply workgroup exceeds LDS");
            op->act().threadsOf[c] = threads;
        }
        return LSG_OK;
    }
    // autotune: for each class set, each size class's workgroup size by timing the apply
    // stage over the admissible multiples of 64 threads (points per thread follow from
    // it); then the set whose tuned apply stage is fastest.  Results do not depend on
    // these choices.
    const int forced = envInt("LSG_CLASS_SET", -1);
    float bestSet = 1e30f;
    int bestCs = op->cs;
    for (int set = 0; set < 3; ++set) {
        if (op->sets[set].nClasses == 0) continue;
        if (forced >= 0 && forced != set) continue;
        op->cs = set;
        auto& A = op->act();
        for (int c = 0; c < A.nClasses; ++c) A.threadsOf[c] = 512;
        for (int c = 0; c < A.nClasses; ++c) {
            float best = 1e30f;
            int bestT = 0;
            for (int t : cands) {
                if (!threadsFit(op, c, t)) continue;
                if (t > 64 && classMax(op, c) <= t / 2) continue;   // half the block covers the largest box
                A.threadsOf[c] = t;
                float tot = 0.f;
                if (int rc = timeApply(op, tot)) return rc;
                if (tot < best) { best = tot; bestT = t; }
            }
            if (!bestT) return fail(LSG_ERR_UNSUPPORTED, "largest depth box does not fit in LDS");
            A.threadsOf[c] = bestT;
        }
        float tot = 0.f;
        if (int rc = timeApply(op, tot)) return rc;
        if (tot < bestSet) { bestSet = tot; bestCs = set; }
    }
    op->cs = bestCs;
    return LSG_OK;
}

int ensureThreads(lsg_op* op) {
    for (int c = 0; c < op->act().nClasses; ++c)
        if (!op->act().threadsOf[c]) return chooseThreads(op, 0);
    return LSG_OK;
}

// host slot-order <-> device order (instance-major batch)
template <typename V>
int uploadSlot(lsg_op* op, const double* h, V* dptr) {
    const Grid& g = *op->G;
    std::vector<V> buf((size_t)g.n * op->batch);
    for (int b = 0; b < op->batch; ++b)
        for (int64_t s = 0; s < g.n; ++s) buf[(size_t)b * g.n + g.devOfSlot[s]] = V(h[(size_t)b * g.n + s]);
    HIPCHK(hipMemcpyAsync(dptr, buf.data(), buf.size() * sizeof(V), hipMemcpyHostToDevice, op->stream));
    HIPCHK(hipStreamSynchronize(op->stream));
    return LSG_OK;
}
template <typename V>
int downloadSlot(lsg_op* op, const V* dptr, double* h) {
    const Grid& g = *op->G;
    std::vector<V> buf((size_t)g.n * op->batch);
    HIPCHK(hipMemcpyAsync(buf.data(), dptr, buf.size() * sizeof(V), hipMemcpyDeviceToHost, op->stream));
    HIPCHK(hipStreamSynchronize(op->stream));
    for (int b = 0; b < op->batch; ++b)
        for (int64_t s = 0; s < g.n; ++s) h[(size_t)b * g.n + s] = double(buf[(size_t)b * g.n + g.devOfSlot[s]]);
    return LSG_OK;
}
int upload(lsg_op* op, const double* h, void* d) {
    return op->prec == 64 ? uploadSlot<double>(op, h, (double*)d) : uploadSlot<float>(op, h, (float*)d);
}
int download(lsg_op* op, const void* d, double* h) {
    return op->prec == 64 ? downloadSlot<double>(op, (const double*)d, h) : downloadSlot<float>(op, (const float*)d, h);
}

int ensureSolver(lsg_op* op) {
    if (op->solverReady) return LSG_OK;
    const size_t nb = (size_t)op->n() * op->batch * op->vs;
    HIPCHK(op->r.alloc(nb));
    HIPCHK(op->h.alloc(nb));
    HIPCHK(op->dv.alloc(nb));
    HIPCHK(op->g.alloc(nb));
    HIPCHK(op->D.alloc(nb));
    HIPCHK(op->f.alloc((size_t)op->n() * op->batch * 8));
    HIPCHK(op->x.alloc((size_t)op->n() * op->batch * 8));
    HIPCHK(op->tmp.alloc((size_t)op->n() * op->batch * 8));
    const int nbk = cg_partial_blocks(op->n());
    HIPCHK(op->partials.alloc((size_t)op->batch * 3 * nbk * sizeof(double)));
    HIPCHK(op->st.alloc((size_t)op->batch * sizeof(CGState)));
    HIPCHK(op->actMap.alloc((size_t)op->batch * sizeof(int)));
    HIPCHK(op->epsInst.alloc((size_t)op->batch * sizeof(double)));
    HIPCHK(op->rows.alloc((size_t)op->batch * sizeof(int)));
    HIPCHK(op->hostSt.alloc((size_t)op->batch * sizeof(CGState)));
    HIPCHK(op->hostMap.alloc((size_t)op->batch * sizeof(int)));
    HIPCHK(op->hostPart.alloc((size_t)op->batch * 3 * nbk * sizeof(double)));
    op->solverReady = true;
    return LSG_OK;
}

// CG chunks replay a captured hipGraph unless LSG_NO_GRAPH is set (direct launches;
// read at every solve)
bool useGraphs() { return std::getenv("LSG_NO_GRAPH") == nullptr; }

// precond3 diagonal of every instance into D (V values, device order)
int launchDiag(lsg_op* op, void* Dp) {
    const Grid& g = *op->G;
    int64_t cs = (int64_t)(op->nA + op->nC) * g.d * op->tabLen;
    if (op->sparse) {
        if (int rc = ensureCellCoef(op)) return rc;
        const int64_t ms = op->nCells * op->nEnt;
        if (op->prec == 64)
            launch_sp_diag<double>(op->sd, op->activeDev.as<uint8_t>(), (int)g.n, op->cellCoef.as<double>(),
                                   2 * op->nCells, op->matKind ? op->cellMat.as<double>() : nullptr, ms, op->nCells,
                                   (double*)Dp, g.n, op->batch, op->stream);
        else
            launch_sp_diag<float>(op->sd, op->activeDev.as<uint8_t>(), (int)g.n, op->cellCoef.as<float>(),
                                  2 * op->nCells, op->matKind ? op->cellMat.as<float>() : nullptr, ms, op->nCells,
                                  (float*)Dp, g.n, op->batch, op->stream);
        HIPCHK(hipGetLastError());
        return LSG_OK;
    }
    if (op->prec == 64)
        launch_diag<double>(op->dSumItems.as<SumItem>(), (int)op->P->sum.size(), g.d, op->nA, op->nC,
                            op->coefTab.as<double>(), cs, op->tabLen, (double*)Dp, g.n, op->batch, op->stream);
    else
        launch_diag<float>(op->dSumItems.as<SumItem>(), (int)op->P->sum.size(), g.d, op->nA, op->nC,
                           op->coefTab.as<float>(), cs, op->tabLen, (float*)Dp, g.n, op->batch, op->stream);
    HIPCHK(hipGetLastError());
    return LSG_OK;
}

int computeDiag(lsg_op* op) {
    if (op->diagValid) return LSG_OK;
    if (int rc = launchDiag(op, op->D.p)) return rc;
    op->diagValid = true;
    return LSG_OK;
}

// PCG on device vectors: f (V), x (V, or double when x64).  Mirrors
// CG::solveHomogen (applications/cg_method.h:759-886).  `open`, if given, lists the
// instances that may need iterations (the others have a zero right-hand side): the
// matvecs launch only those from the first chunk on.
// Work buffers of one CG solve: the op's own (compaction of finished instances and
// chunk graphs allowed), or caller-provided scratch (neither: plain launches only).
struct CgBufs {
    void *r, *h, *dv, *g, *D;
    double* part;
    CGState* st;
    bool own;
    bool pre = false;     // depth-block preconditioner (own buffers only)
};

// Device tables of the depth-block preconditioner of an operator (built once)
int ensureBlockPre(lsg_op* op) {
    if (op->preItems.p) return LSG_OK;
    BlockPre bp;
    std::string err;
    // dense block inverses staged in LDS: up to 128 points in FP32 (64 KB), 64 in FP64 (32 KB);
    // LSG_PRE_DENSE_MAX lowers the limit (larger blocks then take the fast diagonalisation)
    const int dmax = std::min(op->prec == 32 ? 128 : 64, envInt("LSG_PRE_DENSE_MAX", 128));
    if (int rc = buildBlockPre(*op->G, dmax, bp, err)) return fail(rc, err);
    op->preNblk = (int)bp.blocks.size();
    op->preMaxN = 0;
    // FP32 operators apply the dense blocks of at least PRE_MM_MIN points on the matrix
    // cores; the others (and every dense block of an FP64 operator) in blockpre_kernel
    std::vector<int32_t> mm;
    op->preMmMax = 0;
    for (size_t b = 0; b < bp.blocks.size(); ++b) {
        const PreBlock& pb = bp.blocks[b];
        if (pb.mode != PRE_DENSE) continue;
        if (op->prec == 32 && pb.n >= PRE_MM_MIN && !envInt("LSG_NO_PRE_MM", 0)) {
            mm.push_back((int32_t)b);
            op->preMmMax = std::max(op->preMmMax, (int)pb.n);
        } else {
            op->preMaxN = std::max(op->preMaxN, (int)pb.n);
        }
    }
    op->preNmm = (int)mm.size();
    HIPCHK(op->preItems.alloc(bp.blocks.size() * sizeof(PreBlock)));
    HIPCHK(hipMemcpy(op->preItems.p, bp.blocks.data(), bp.blocks.size() * sizeof(PreBlock), hipMemcpyHostToDevice));
    auto upload = [&](DevMem& m, const std::vector<double>& v) -> int {
        HIPCHK(m.alloc(std::max<size_t>(1, v.size()) * op->vs));
        if (v.empty()) return LSG_OK;
        if (op->prec == 64) {
            HIPCHK(hipMemcpy(m.p, v.data(), v.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(v.begin(), v.end());
            HIPCHK(hipMemcpy(m.p, f.data(), f.size() * 4, hipMemcpyHostToDevice));
        }
        return LSG_OK;
    };
    if (int rc = upload(op->preInv, bp.inv)) return rc;
    if (int rc = upload(op->preD0, bp.d0)) return rc;
    if (int rc = upload(op->preFdV, bp.fdV)) return rc;
    if (int rc = upload(op->preFdLam, bp.fdLam)) return rc;
    op->preNfd = (int)bp.fdList.size();
    op->preMaxFd = bp.maxFd;
    auto upload32 = [&](DevMem& m, const std::vector<int32_t>& v) -> int {
        HIPCHK(m.alloc(std::max<size_t>(1, v.size()) * sizeof(int32_t)));
        if (!v.empty()) HIPCHK(hipMemcpy(m.p, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        return LSG_OK;
    };
    if (int rc = upload32(op->preFdList, bp.fdList)) return rc;
    if (int rc = upload32(op->preFdPerm, bp.fdPerm)) return rc;
    if (int rc = upload32(op->preMmList, mm)) return rc;
    HIPCHK(op->prePart.alloc((size_t)op->batch * op->preNblk * sizeof(double)));
    return LSG_OK;
}

template <typename V>
BlockPreArgs<V> blockPreArgs(lsg_op* op, bool on) {
    BlockPreArgs<V> a;
    if (!on) return a;
    a.items = op->preItems.as<PreBlock>();
    a.nblk = op->preNblk;
    a.maxN = op->preMaxN;
    a.inv = op->preInv.as<V>();
    a.d0 = op->preD0.as<V>();
    a.part = op->prePart.as<double>();
    a.d = op->G->d;
    a.nfd = op->preNfd;
    a.maxFd = op->preMaxFd;
    a.fdList = op->preFdList.as<int32_t>();
    a.fdPerm = op->preFdPerm.as<int32_t>();
    a.nmm = op->preNmm;
    a.mmMax = op->preMmMax;
    a.mmList = op->preMmList.as<int32_t>();
    a.mmGroups = envInt("LSG_PRE_MM_GROUPS", 1);
    a.fdV = op->preFdV.as<V>();
    a.fdLam = op->preFdLam.as<V>();
    return a;
}

template <typename V>
int cgRun(lsg_op* op, const CgBufs& bf, const V* f, void* x, bool x64, double eps, int rule, int maxit,
          std::vector<CGState>& out, bool xZero, const double* epsInst, const std::vector<int>* open);

template <typename V>
int cgDevice(lsg_op* op, const V* f, void* x, bool x64, double eps, int rule, int maxit,
             std::vector<CGState>& out, bool xZero = false, const double* epsInst = nullptr,
             const std::vector<int>* open = nullptr, bool precond = true) {
    const int64_t n = op->n();
    const int B = op->batch;
    int rc = ensureSolver(op);
    if (rc) return rc;
    // the cell coefficients are refreshed eagerly here, outside any captured chunk: a chunk
    // replayed after lsg_op_set_coef then reads the current ones
    if ((rc = ensureCellCoef(op))) return rc;
    if (precond) {
        if ((rc = computeDiag(op))) return rc;
    } else if (!op->ones.p) {     // unpreconditioned CG: D = 1
        HIPCHK(op->ones.alloc((size_t)n * B * op->vs));
        if (op->prec == 64) {
            std::vector<double> one((size_t)n * B, 1.0);
            HIPCHK(hipMemcpy(op->ones.p, one.data(), one.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> one((size_t)n * B, 1.0f);
            HIPCHK(hipMemcpy(op->ones.p, one.data(), one.size() * 4, hipMemcpyHostToDevice));
        }
    }
    if ((rc = ensureThreads(op))) return rc;
    CgBufs bf{op->r.p, op->h.p, op->dv.p, op->g.p, precond ? op->D.p : op->ones.p, op->partials.as<double>(),
              op->st.as<CGState>(), true};
    if (precond && op->precKind == LSG_PRECOND_DEPTH_BLOCK) {
        if ((rc = ensureBlockPre(op))) return rc;
        bf.pre = true;
    }
    return cgRun<V>(op, bf, f, x, x64, eps, rule, maxit, out, xZero, epsInst, open);
}

// PCG (CG::solveHomogen, applications/cg_method.h:759-886) on the buffers bf, in phases
// that only issue work on the operator's stream (cgStart, cgIssueChunk) or read back
// what a completed chunk left (cgCollectChunk, after the stream has been synchronised),
// so that one host thread can advance the solves of several operators at once.
template <typename V>
struct CgPhase {
    lsg_op* op = nullptr;
    CgBufs bf{};
    const V* f = nullptr;
    void* x = nullptr;
    bool x64 = false;
    int maxit = 0;
    int done = 0;
    int lastK = 0;
    bool finished = false;
    std::vector<CGState> out;
};

// host-visible copies of the CG state: pinned memory of the operator (asynchronous copies
// from several streams do not serialise the host), or, without own buffers, out itself
template <typename V>
CGState* cgHostState(CgPhase<V>& c) {
    return c.bf.own ? c.op->hostSt.template as<CGState>() : c.out.data();
}

template <typename V>
int cgStart(CgPhase<V>& c, double eps, int rule, bool xZero, const double* epsInst, const std::vector<int>* open) {
    lsg_op* op = c.op;
    const int64_t n = op->n();
    const int B = op->batch;
    hipStream_t s = op->stream;
    const int nb = cg_partial_blocks(n);
    V* r = (V*)c.bf.r;
    V* g = (V*)c.bf.g;
    // r = A x (x in V for the matvec); a start vector known to be zero needs no matvec
    if (xZero) {
        launch_zero(r, n * B * (int64_t)sizeof(V), s);
    } else {
        const V* xv;
        if (c.x64 && sizeof(V) != 8) {
            launch_cast_d2f((const double*)c.x, (float*)g, n * B, s);
            xv = g;
        } else xv = (const V*)c.x;
        if (int rc = runMatvec<V>(op, xv, r, nullptr)) return rc;
    }
    const BlockPreArgs<V> pre = blockPreArgs<V>(op, c.bf.pre);
    launch_cg_init<V>(c.f, r, (V*)c.bf.h, (V*)c.bf.dv, (V*)c.bf.D, n, B, c.bf.part, nb, s, pre);
    launch_cg_start(c.bf.st, c.bf.part, nb, eps, epsInst, rule, c.maxit, B, s, pre.part, pre.nblk);
    HIPCHK(hipGetLastError());
    c.out.resize(B);
    c.done = 0;
    c.finished = c.maxit <= 0;
    op->nMap = 0;
    if (c.bf.own && open && !open->empty() && (int)open->size() < B) {
        int* hm = op->hostMap.as<int>();
        std::copy(open->begin(), open->end(), hm);
        HIPCHK(hipMemcpyAsync(op->actMap.p, hm, open->size() * sizeof(int), hipMemcpyHostToDevice, s));
        op->nMap = (int)open->size();
    }
    return LSG_OK;
}

// the next chunk of up to 8 iterations, then the state copied towards the host
template <typename V>
int cgIssueChunk(CgPhase<V>& c) {
    lsg_op* op = c.op;
    const int64_t n = op->n();
    const int B = op->batch;
    hipStream_t s = op->stream;
    const int nb = cg_partial_blocks(n);
    V* r = (V*)c.bf.r;
    V* h = (V*)c.bf.h;
    V* dv = (V*)c.bf.dv;
    V* g = (V*)c.bf.g;
    V* D = (V*)c.bf.D;
    double* part = c.bf.part;
    CGState* st = c.bf.st;
    void* x = c.x;
    const bool x64 = c.x64;
    int rc = LSG_OK;
    const BlockPreArgs<V> pre = blockPreArgs<V>(op, c.bf.pre);
    const int k = std::min(8, c.maxit - c.done);
    if (c.bf.own && op->nMap == 0 && useGraphs()) {
        // the chunk's launches depend only on the operator, x and k: capture once, then
        // replay (the kernels read the convergence state on the device)
        if (!op->chunkExec || op->chunkX != x || op->chunkK != k || op->chunkX64 != x64 || op->chunkD != D ||
            op->chunkPre != c.bf.pre) {
            if (op->chunkExec) HIPCHK(hipGraphExecDestroy(op->chunkExec));
            op->chunkExec = nullptr;
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            op->capturing = true;
            for (int q = 0; q < k && !rc; ++q) {
                rc = runMatvec<V>(op, dv, g, st);
                if (!rc) launch_cg_iter_tail<V>(r, x, x64, h, dv, g, D, n, B, part, nb, st, s, pre);
            }
            op->capturing = false;
            const hipError_t ec = hipStreamEndCapture(s, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            HIPCHK(ec);
            const hipError_t ei = hipGraphInstantiate(&op->chunkExec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(ei);
            op->chunkX = x;
            op->chunkK = k;
            op->chunkX64 = x64;
            op->chunkD = D;
            op->chunkPre = c.bf.pre;
        }
        HIPCHK(hipGraphLaunch(op->chunkExec, s));
    } else {
        for (int q = 0; q < k; ++q) {
            if ((rc = runMatvec<V>(op, dv, g, st))) return rc;
            launch_cg_iter_tail<V>(r, x, x64, h, dv, g, D, n, B, part, nb, st, s, pre);
        }
    }
    c.lastK = k;
    HIPCHK(hipMemcpyAsync(cgHostState(c), st, sizeof(CGState) * B, hipMemcpyDeviceToHost, s));
    return LSG_OK;
}

// after the stream is synchronised: the state of the chunk just run; finished
// instances drop out of the matvec grids (their workgroups would only read their flag
// and exit, and they delay the dispatch of the ones still running)
template <typename V>
int cgCollectChunk(CgPhase<V>& c) {
    lsg_op* op = c.op;
    const int B = op->batch;
    const CGState* hs = cgHostState(c);
    if (hs != c.out.data()) std::copy(hs, hs + B, c.out.begin());
    c.done += c.lastK;
    int nact = 0;
    int* hm = c.bf.own ? op->hostMap.as<int>() : nullptr;
    for (int b = 0; b < B; ++b)
        if (c.out[b].active) {
            if (hm) hm[nact] = b;
            ++nact;
        }
    if (nact == 0 || c.done >= c.maxit) {
        c.finished = true;
        op->nMap = 0;
        return LSG_OK;
    }
    if (c.bf.own && nact < B) {
        HIPCHK(hipMemcpyAsync(op->actMap.p, hm, nact * sizeof(int), hipMemcpyHostToDevice, op->stream));
        op->nMap = nact;
    }
    return LSG_OK;
}

template <typename V>
int cgRun(lsg_op* op, const CgBufs& bf, const V* f, void* x, bool x64, double eps, int rule, int maxit,
          std::vector<CGState>& out, bool xZero, const double* epsInst, const std::vector<int>* open) {
    CgPhase<V> c;
    c.op = op;
    c.bf = bf;
    c.f = f;
    c.x = x;
    c.x64 = x64;
    c.maxit = maxit;
    int rc = cgStart<V>(c, eps, rule, xZero, epsInst, open);
    if (rc) return rc;
    if (c.finished) {      // no iteration allowed: the start state only
        HIPCHK(hipMemcpyAsync(c.out.data(), bf.st, sizeof(CGState) * op->batch, hipMemcpyDeviceToHost, op->stream));
    }
    while (!c.finished) {
        if ((rc = cgIssueChunk<V>(c))) return rc;
        HIPCHK(hipStreamSynchronize(op->stream));
        if ((rc = cgCollectChunk<V>(c))) return rc;
    }
    HIPCHK(hipStreamSynchronize(op->stream));
    out = c.out;
    return LSG_OK;
}

// True relative residual ||f - A x|| / ||f|| in FP64 (x, f double, device), issued on
// op64's stream; residualCollect reads it after the stream is synchronised.  With a
// non-null `sub` (a strict subset of the instances), only those instances' r and rel
// are recomputed; the others keep theirs.
int residualIssue(lsg_op* op64, const double* f, const double* x, double* tmpAx, double* r,
                  const std::vector<int>* sub) {
    const int64_t n = op64->n();
    const int B = op64->batch;
    hipStream_t s = op64->stream;
    const int nSub = (sub && (int)sub->size() < B) ? (int)sub->size() : 0;
    if (nSub > 0) {
        int* hm = op64->hostMap.as<int>();
        std::copy(sub->begin(), sub->end(), hm);
        HIPCHK(hipMemcpyAsync(op64->actMap.p, hm, nSub * sizeof(int), hipMemcpyHostToDevice, s));
    }
    int rc = runMatvec<double>(op64, x, tmpAx, nullptr, nSub);
    if (rc) return rc;
    const int nb = cg_partial_blocks(n);
    launch_residual(f, tmpAx, r, n, nSub > 0 ? nSub : B, op64->partials.as<double>(), nb, s,
                    nSub > 0 ? op64->actMap.as<int>() : nullptr);
    HIPCHK(hipMemcpyAsync(op64->hostPart.p, op64->partials.p, (size_t)B * 2 * nb * 8, hipMemcpyDeviceToHost, s));
    return LSG_OK;
}

void residualCollect(lsg_op* op64, std::vector<double>& rel, const std::vector<int>* sub) {
    const int64_t n = op64->n();
    const int B = op64->batch;
    const int nb = cg_partial_blocks(n);
    const double* hp = op64->hostPart.as<double>();
    const int nSub = (sub && (int)sub->size() < B) ? (int)sub->size() : 0;
    rel.resize(B);
    auto one = [&](int b) {
        double rr = 0, ff = 0;
        for (int k = 0; k < nb; ++k) rr += hp[((size_t)b * 2 + 0) * nb + k];
        for (int k = 0; k < nb; ++k) ff += hp[((size_t)b * 2 + 1) * nb + k];
        rel[b] = ff > 0 ? std::sqrt(rr / ff) : std::sqrt(rr);
    };
    if (nSub > 0)
        for (int b : *sub) one(b);
    else
        for (int b = 0; b < B; ++b) one(b);
}

int trueResidual(lsg_op* op64, const double* f, const double* x, double* tmpAx, double* r, std::vector<double>& rel,
                 const std::vector<int>* sub = nullptr) {
    if (int rc = residualIssue(op64, f, x, tmpAx, r, sub)) return rc;
    HIPCHK(hipStreamSynchronize(op64->stream));
    residualCollect(op64, rel, sub);
    return LSG_OK;
}

// ---- mixed-precision solves of several operator pairs at once (lsg_ir_solve_many)
// One problem group: an FP64 / FP32 operator pair on one grid with its batch.  Both
// operators of a group issue on one stream (the FP32 operator's) for the duration of the
// solve, so a group's phases need no events between them.
struct IrGroup {
    lsg_op* op64 = nullptr;
    lsg_op* op32 = nullptr;
    double* f = nullptr;
    double* x = nullptr;
    int B = 0;
    int64_t n = 0;
    int sweep = 0;
    enum Phase { RESIDUAL, RESTART, INNER, DONE } phase = RESIDUAL;
    std::vector<int> iters, outer, conv, newly, open;
    std::vector<double> rel, epsH;
    CgPhase<float> cg;
    hipStream_t saved64 = nullptr;
};

namespace {
// the FP64 residual of the group's open instances (issued on the group's stream)
int irIssueResidual(IrGroup& g) {
    g.open.clear();
    for (int b = 0; b < g.B; ++b)
        if (!g.conv[b]) g.open.push_back(b);
    return residualIssue(g.op64, g.f, g.x, g.op64->tmp.as<double>(), g.op64->r.as<double>(), &g.open);
}

// after the residual: converged instances, inner tolerances; then either the end of the
// group's solve or its next inner FP32 PCG (right-hand side, start, first chunk)
int irAfterResidual(IrGroup& g, double rel_tol, double inner_rel_tol, int maxit) {
    residualCollect(g.op64, g.rel, g.phase == IrGroup::RESTART ? nullptr : &g.open);
    if (g.sweep == 0 && g.phase == IrGroup::RESIDUAL) {
        // a start vector worse than zero (||f - A x0|| > ||f||) is replaced by zero
        bool redo = false;
        for (int b = 0; b < g.B; ++b)
            if (g.rel[b] > 1.0) {
                HIPCHK(hipMemsetAsync(g.x + (size_t)b * g.n, 0, g.n * 8, g.op64->stream));
                redo = true;
            }
        if (redo) {
            g.phase = IrGroup::RESTART;
            return residualIssue(g.op64, g.f, g.x, g.op64->tmp.as<double>(), g.op64->r.as<double>(), nullptr);
        }
    }
    bool all = true;
    g.newly.clear();
    for (int b = 0; b < g.B; ++b) {
        if (g.rel[b] != g.rel[b]) return fail(LSG_ERR_NAN, "NaN residual");
        if (g.rel[b] <= rel_tol) {
            if (!g.conv[b]) g.newly.push_back(b);
            g.conv[b] = 1;
        } else all = false;
        // inner tolerance: inner_rel_tol, or less when the remaining reduction is smaller.
        // The first sweep aims at inner_rel_tol * ||f|| rather than at inner_rel_tol times
        // the start residual, so that a good start (rel < 1) shortens it; from a zero
        // start (rel = 1) both are the same.
        const double rr = std::max(g.rel[b], 1e-300);
        const double target = g.sweep == 0 ? inner_rel_tol / std::min(rr, 1.0) : inner_rel_tol;
        g.epsH[b] = std::min(0.5, std::max(target, 0.3 * rel_tol / rr));
    }
    if (all || g.sweep == maxit) {
        g.phase = IrGroup::DONE;
        return LSG_OK;
    }
    // inner right-hand side r32 = f - A x in FP32, rows of converged instances zero
    hipStream_t s = g.op32->stream;
    const int64_t tot = g.n * g.B;
    double* r64 = g.op64->r.as<double>();
    if (!g.newly.empty()) {
        int* hm = g.op64->hostMap.as<int>();
        std::copy(g.newly.begin(), g.newly.end(), hm);
        HIPCHK(hipMemcpyAsync(g.op64->rows.p, hm, g.newly.size() * sizeof(int), hipMemcpyHostToDevice, s));
        launch_zero_rows(r64, g.n, g.op64->rows.as<int>(), (int)g.newly.size(), s);
    }
    double* he = g.op64->hostPart.as<double>();      // (the residual partials are read already)
    std::copy(g.epsH.begin(), g.epsH.end(), he);
    HIPCHK(hipMemcpyAsync(g.op64->epsInst.p, he, (size_t)g.B * 8, hipMemcpyHostToDevice, s));
    launch_cast_d2f(r64, g.op32->f.as<float>(), tot, s);
    g.open.clear();
    for (int b = 0; b < g.B; ++b)
        if (!g.conv[b]) g.open.push_back(b);
    // inner PCG from e = 0
    CgPhase<float>& c = g.cg;
    c = CgPhase<float>();
    c.op = g.op32;
    int rc = computeDiag(g.op32);
    if (rc) return rc;
    c.bf = CgBufs{g.op32->r.p, g.op32->h.p, g.op32->dv.p, g.op32->g.p, g.op32->D.p, g.op32->partials.as<double>(),
                  g.op32->st.as<CGState>(), true};
    if (g.op32->precKind == LSG_PRECOND_DEPTH_BLOCK) {
        if ((rc = ensureBlockPre(g.op32))) return rc;
        c.bf.pre = true;
    }
    c.f = g.op32->f.as<float>();
    c.x = g.op32->x.p;
    c.x64 = false;
    c.maxit = 1000;
    launch_zero(c.x, tot * 4, s);
    if ((rc = cgStart<float>(c, inner_rel_tol, LSG_STOP_REL, true, g.op64->epsInst.as<double>(), &g.open))) return rc;
    g.phase = IrGroup::INNER;
    return cgIssueChunk<float>(c);
}

// after an inner chunk: the next chunk, or the correction x += e and the next residual
int irAfterChunk(IrGroup& g) {
    int rc = cgCollectChunk<float>(g.cg);
    if (rc) return rc;
    if (!g.cg.finished) return cgIssueChunk<float>(g.cg);
    launch_axpy_correction(g.x, g.op32->x.as<float>(), g.n * g.B, g.op32->stream);
    for (int b = 0; b < g.B; ++b)
        if (!g.conv[b]) {
            g.iters[b] += g.cg.out[b].it;
            g.outer[b] += 1;
        }
    ++g.sweep;
    g.phase = IrGroup::RESIDUAL;
    return irIssueResidual(g);
}
}  // namespace

// FP32 inner PCG under FP64 iterative refinement for every group.  One host thread drives
// a state machine per group (residual -> inner PCG chunks -> correction -> residual ...):
// it polls the groups' streams and advances each group as soon as its last phase has
// completed, so small grids finish their sweeps while a large one is still iterating, and
// the GPU runs the groups side by side.
int irSolveGroups(std::vector<IrGroup>& G, double rel_tol, double inner_rel_tol, int maxit) {
    int rc = LSG_OK;
    for (IrGroup& g : G) {
        if ((rc = ensureSolver(g.op64)) || (rc = ensureSolver(g.op32)) || (rc = ensureThreads(g.op64)) ||
            (rc = ensureThreads(g.op32)) || (rc = ensureCellCoef(g.op32)) || (rc = ensureCellCoef(g.op64)))
            return rc;
        HIPCHK(hipStreamSynchronize(g.op64->stream));
        g.B = g.op64->batch;
        g.n = g.op64->n();
        g.iters.assign(g.B, 0);
        g.outer.assign(g.B, 0);
        g.conv.assign(g.B, 0);
        g.epsH.assign(g.B, 0.0);
        g.sweep = 0;
        g.phase = IrGroup::RESIDUAL;
    }
    for (IrGroup& g : G) {          // the group's FP64 work issues on its FP32 operator's stream
        g.saved64 = g.op64->stream;
        g.op64->stream = g.op32->stream;
    }
    for (IrGroup& g : G)
        if ((rc = irIssueResidual(g))) break;
    while (rc == LSG_OK) {
        bool busy = false, advanced = false;
        for (IrGroup& g : G) {
            if (g.phase == IrGroup::DONE) continue;
            busy = true;
            const hipError_t q = hipStreamQuery(g.op32->stream);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) {
                rc = fail(LSG_ERR_DEVICE, std::string("hipStreamQuery: ") + hipGetErrorString(q));
                break;
            }
            advanced = true;
            rc = g.phase == IrGroup::INNER ? irAfterChunk(g) : irAfterResidual(g, rel_tol, inner_rel_tol, maxit);
            if (rc) break;
        }
        if (!busy || rc) break;
        if (!advanced) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (IrGroup& g : G) {
        (void)hipStreamSynchronize(g.op32->stream);
        g.op64->stream = g.saved64;
        g.op32->nMap = 0;
        g.op64->nMap = 0;
    }
    return rc;
}

// Tables and work vectors of the basis transformations (FP64 operators).
int ensureTransform(lsg_op* op) {
    if (op->tr) return LSG_OK;
    if (op->prec != 64) return fail(LSG_ERR_UNSUPPORTED, "basis transformations run on an FP64 operator");
    const Grid& g = op->grid->g;
    auto tr = std::make_unique<TrDev>();
    std::string err;
    if (int rc = buildTransform(g, tr->host, err)) return fail(rc, err);
    const TrTables& h = tr->host;
    if (!h.adaptive && h.maxBox * 8 > 65536)
        return fail(LSG_ERR_UNSUPPORTED, "depth box too large for the transform workgroup");
    auto up = [&](DevMem& m, const void* src, size_t bytes) -> hipError_t {
        hipError_t e = m.alloc(bytes);
        if (e != hipSuccess) return e;
        return bytes ? hipMemcpy(m.p, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
    };
    HIPCHK(up(tr->items, h.items.data(), h.items.size() * sizeof(TrItem)));
    HIPCHK(up(tr->gidx, h.gidx.data(), h.gidx.size() * 4));
    HIPCHK(up(tr->list, h.levelList.data(), h.levelList.size() * 4));
    HIPCHK(up(tr->rowDev, h.rowDev.data(), h.rowDev.size() * 4));
    HIPCHK(up(tr->rowPtr, h.rowPtr.data(), h.rowPtr.size() * 8));
    HIPCHK(up(tr->col, h.col.data(), h.col.size() * 4));
    HIPCHK(up(tr->fac, h.fac.data(), h.fac.size() * 8));
    HIPCHK(up(tr->poles, h.poles.data(), h.poles.size() * 4));
    HIPCHK(up(tr->byLevel, h.byLevel.data(), h.byLevel.size() * 4));
    HIPCHK(up(tr->bstat, h.bstat.data(), h.bstat.size()));
    HIPCHK(up(tr->comps, h.comps.data(), h.comps.size() * sizeof(TrComp)));
    HIPCHK(up(tr->compOff, h.compOff.data(), h.compOff.size() * 4));
    HIPCHK(up(tr->afac, h.afac.data(), h.afac.size() * 8));
    const size_t nb = (size_t)g.n * op->batch * 8;
    HIPCHK(tr->hier.alloc(nb));
    HIPCHK(tr->work.alloc(nb));
    HIPCHK(tr->work2.alloc(nb));
    HIPCHK(tr->scratch.alloc((size_t)h.boxTot * op->batch * 8));
    HIPCHK(tr->norms.alloc((size_t)op->batch * 8));
    op->tr = std::move(tr);
    return LSG_OK;
}

enum { TR_LEVEL = 2, TR_SYNTH = 3 };   // tr_box_kernel modes

void trBox(lsg_op* op, int mode, const int32_t* list, int count, int64_t maxBox, const double* in, double* out) {
    const Grid& g = op->grid->g;
    TrDev& t = *op->tr;
    launch_tr_box(mode, t.items.as<TrItem>(), list, count, t.gidx.as<int32_t>(), t.fac.as<double>(), t.host.facTot,
                  maxBox, in, out, t.scratch.as<double>(), g.n, t.host.boxTot, op->batch, op->stream);
}

void trGather(lsg_op* op, size_t group, double* v, int sub) {
    const Grid& g = op->grid->g;
    TrDev& t = *op->tr;
    const auto& gr = t.host.groups[group];
    launch_tr_gather(t.rowDev.as<int32_t>() + gr[0], t.rowPtr.as<int64_t>() + gr[0], t.col.as<int32_t>(), gr[1],
                     t.scratch.as<double>(), v, g.n, t.host.boxTot, sub, op->batch, op->stream);
}

// CalcHierarchicalBasis as d point-parallel passes (one per dimension, in dimension
// order), ping-ponging between out and the scratch so that the last pass writes out
void trHierPoles(lsg_op* op, const double* in, double* out) {
    const Grid& g = op->grid->g;
    TrDev& t = *op->tr;
    const int d = g.d;
    const double* src = in;
    for (int i = 0; i < d; ++i) {
        double* dst = ((d - 1 - i) % 2 == 0) ? out : t.scratch.as<double>();
        launch_tr_pole(t.poles.as<int32_t>() + (size_t)i * g.n * 2, src, dst, g.n, op->batch, op->stream);
        src = dst;
    }
}

// CalcNodalBasis in place on v: per dimension (in dimension order), one launch per depth
// along that dimension, coarse to fine
int trDehierPoles(lsg_op* op, double* v) {
    const Grid& g = op->grid->g;
    TrDev& t = *op->tr;
    for (int i = 0; i < g.d; ++i) {
        const int64_t* off = &t.host.lvlOff[(size_t)i * (MAXLEV + 2)];
        for (int l = 2; l <= MAXLEV; ++l)
            launch_tr_dehier(t.byLevel.as<int32_t>() + (size_t)i * g.n + off[l], off[l + 1] - off[l],
                             t.poles.as<int32_t>() + (size_t)i * g.n * 2, v, g.n, op->batch, op->stream);
    }
    HIPCHK(hipGetLastError());
    return LSG_OK;
}

// One adaptive-grid box launch (mode 0: depth `depth` of calcPrewByNodal, 1: every depth's
// prewavelets in hierarchical form into the scratch).
void trAdapt(lsg_op* op, int mode, int depth, const double* src, double* hier, double* out, int first) {
    TrDev& t = *op->tr;
    const TrTables& h = t.host;
    const int nitems = mode == 0 ? 1 : (int)h.items.size();
    const int64_t mb = mode == 0 ? h.items[depth].total : h.maxBox;
    launch_tr_adapt(mode, t.items.as<TrItem>(), depth, nitems, mb, t.comps.as<TrComp>(), t.compOff.as<int32_t>(),
                    t.gidx.as<int32_t>(), t.bstat.as<uint8_t>(), t.afac.as<double>(), h.afacTot, src, hier, out,
                    t.scratch.as<double>(), op->n(), h.boxTot, first, trLdsLimit(op->maxLds), op->batch,
                    op->stream);
}

// Device vectors (all instances, device order): what 0 hierarchize, 1 dehierarchize,
// 2 prewavelets from nodal values, 3 nodal values from prewavelets.
int transformDev(lsg_op* op, int what, const double* in, double* out) {
    int rc = ensureTransform(op);
    if (rc) return rc;
    TrDev& t = *op->tr;
    const int nd = (int)t.host.items.size();
    const int64_t mb = t.host.maxBox;
    if (t.host.adaptive && what == 2) {   // calcPrewByNodal, depth by depth (BasisTransformations.cpp:486-545)
        double* hier = t.hier.as<double>();
        trHierPoles(op, in, hier);
        HIPCHK(hipMemsetAsync(out, 0, (size_t)op->n() * op->batch * 8, op->stream));
        for (size_t k = 0; k < t.host.seq.size(); ++k)
            trAdapt(op, 0, t.host.seq[k], k == 0 ? in : hier, hier, out, k == 0);
        HIPCHK(hipGetLastError());
        return LSG_OK;
    }
    if (t.host.adaptive && what == 3) {   // calcNodalByPrew (BasisTransformations.cpp:670-707)
        double* hier = t.hier.as<double>();
        trAdapt(op, 1, -1, in, nullptr, nullptr, 0);
        // the gather writes the active points only; hanging points start (and stay) at 0
        HIPCHK(hipMemsetAsync(hier, 0, (size_t)op->n() * op->batch * 8, op->stream));
        trGather(op, t.host.groups.size() - 1, hier, 0);
        HIPCHK(hipMemcpyAsync(out, hier, (size_t)op->n() * op->batch * 8, hipMemcpyDeviceToDevice, op->stream));
        if (int rc2 = trDehierPoles(op, out)) return rc2;
        HIPCHK(hipGetLastError());
        return LSG_OK;
    }
    switch (what) {
        case 0: trHierPoles(op, in, out); break;
        case 1:
            HIPCHK(hipMemcpyAsync(out, in, (size_t)op->n() * op->batch * 8, hipMemcpyDeviceToDevice, op->stream));
            if (int rc2 = trDehierPoles(op, out)) return rc2;
            break;
        case 2: {   // calcPrewByNodal (BasisTransformations.cpp:486-545)
            double* hier = t.hier.as<double>();
            trHierPoles(op, in, hier);
            for (size_t l = 0; l < t.host.levels.size(); ++l) {
                trBox(op, TR_LEVEL, t.list.as<int32_t>() + t.host.levels[l][0], t.host.levels[l][1],
                      t.host.levelMaxBox[l], hier, out);
                if (l + 1 < t.host.levels.size()) trGather(op, l, hier, 1);
            }
            break;
        }
        case 3: {   // calcNodalByPrew (BasisTransformations.cpp:670-707)
            double* hier = t.hier.as<double>();
            trBox(op, TR_SYNTH, nullptr, nd, mb, in, nullptr);
            trGather(op, t.host.groups.size() - 1, hier, 0);
            HIPCHK(hipMemcpyAsync(out, hier, (size_t)op->n() * op->batch * 8, hipMemcpyDeviceToDevice, op->stream));
            if (int rc2 = trDehierPoles(op, out)) return rc2;
            break;
        }
        default: return fail(LSG_ERR_ARG, "unknown transform");
    }
    HIPCHK(hipGetLastError());
    return LSG_OK;
}

int transformAbi(lsg_op* op, int what, const void* in, void* out, int layout) {
    if (!op || !in || !out) return fail(LSG_ERR_ARG, "null argument");
    int rc = ensureTransform(op);
    if (rc) return rc;
    if (layout == LSG_DEVICE) {
        if (in == out) return fail(LSG_ERR_ARG, "transform input and output must differ");
        if ((rc = transformDev(op, what, (const double*)in, (double*)out))) return rc;
        HIPCHK(hipStreamSynchronize(op->stream));
        return LSG_OK;
    }
    if ((rc = upload(op, (const double*)in, op->tr->work.p))) return rc;
    if ((rc = transformDev(op, what, op->tr->work.as<double>(), op->tr->work2.as<double>()))) return rc;
    return download(op, op->tr->work2.p, (double*)out);
}

// ---- CG without device allocation (lsg_cg_solve_noalloc)
// Work buffers of that entry point: the CG vectors r, h, d, g, the diagonal D, the
// partial sums and the per-instance CG state.  One copy per device, held until the
// process ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (noallocScratch), never through the symbol itself.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratchMu;
void* g_scratchDev[256] = {};

// device address of `scratch` on the current device (looked up once per device)
int noallocScratch(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) return fail(LSG_ERR_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchDev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(LSG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchDev[dev] = p;
    }
    *out = g_scratchDev[dev];
    return LSG_OK;
}

template <typename V>
int cgNoalloc(lsg_op* op, const void* f, void* x, double eps, int stopRule, int maxit, std::vector<CGState>& st) {
    const bool precond = !(stopRule & LSG_CG_NO_PRECOND);
    const int rule = stopRule & ~LSG_CG_NO_PRECOND;
    if (rule != LSG_STOP_REF_ABS && rule != LSG_STOP_REL) return fail(LSG_ERR_ARG, "unknown stopping rule");
    const int64_t n = op->n();
    const int B = op->batch;
    const int nb = cg_partial_blocks(n);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t vec = up((size_t)n * B * sizeof(V));
    const size_t part = up((size_t)B * 3 * nb * sizeof(double));
    const size_t stb = up((size_t)B * sizeof(CGState));
    const size_t need = 5 * vec + part + stb;
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(LSG_ERR_CAPACITY, "lsg_cg_solve_noalloc: " + std::to_string(need) + " bytes of work buffers (" +
                                          std::to_string(n) + " points x " + std::to_string(B) +
                                          " instances) exceed its " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                          "-byte scratch; use lsg_cg_solve");
    void* base = nullptr;
    if (int rc = noallocScratch(&base)) return rc;
    unsigned char* p = (unsigned char*)base;
    CgBufs bf{p, p + vec, p + 2 * vec, p + 3 * vec, p + 4 * vec, (double*)(p + 5 * vec),
              (CGState*)(p + 5 * vec + part), false};
    // every scratch buffer is written before it is read (the diagonal and D = 1 here, r by
    // the start matvec, h, d and the partials by the init kernel, the state by the start
    // kernel), so nothing carried over from an earlier call is used
    int rc = ensureCellCoef(op);
    if (rc) return rc;
    if (precond) {
        if ((rc = launchDiag(op, bf.D))) return rc;
    } else {
        launch_fill_one(bf.D, n * B, sizeof(V) == 8, op->stream);
    }
    if ((rc = ensureThreads(op))) return rc;
    return cgRun<V>(op, bf, (const V*)f, x, sizeof(V) == 8, eps, rule, maxit, st, false, nullptr, nullptr);
}

void planStats(const Grid& g, const Plan& P, int64_t* out, int nout) {
    int64_t v[12] = {(int64_t)g.depths.size(), g.nml, P.nApply, (int64_t)P.apply.size(),
                     (int64_t)P.streamValues, P.maxBox, P.caseSize, P.wSize, P.partSize, (int64_t)P.movedValues,
                     (int64_t)P.applyValues, (int64_t)P.streamApplyValues};
    for (int i = 0; i < nout && i < 12; ++i) out[i] = v[i];
}

}  // namespace

extern "C" {

const char* lsg_last_error(void) { return g_err.c_str(); }

int lsg_device_available(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n > 0;
}

int lsg_grid_create(int d, int64_t n, const int32_t* idx, const uint8_t* active, lsg_grid** out) {
    if (!out) return fail(LSG_ERR_ARG, "out is null");
    auto* g = new lsg_grid();
    std::string err;
    int rc = buildGrid(d, n, idx, active, g->g, err);
    if (rc) { delete g; return fail(rc, err); }
    *out = g;
    return LSG_OK;
}

int lsg_grid_regular(int d, int level, lsg_grid** out) {
    if (d < 1 || d > MAXD || level < 0 || level > 14) return fail(LSG_ERR_ARG, "bad dimension or level");
    std::vector<int32_t> idx;
    regularIndexTable(d, level, idx);
    return lsg_grid_create(d, (int64_t)(idx.size() / d), idx.data(), nullptr, out);
}

int64_t lsg_grid_size(const lsg_grid* g) { return g ? g->g.n : -1; }
int lsg_grid_dim(const lsg_grid* g) { return g ? g->g.d : -1; }

int lsg_grid_slot_order(const lsg_grid* g, int32_t* idx_out, uint8_t* active_out) {
    if (!g) return fail(LSG_ERR_ARG, "null grid");
    if (idx_out) std::memcpy(idx_out, g->g.idx.data(), g->g.idx.size() * sizeof(int32_t));
    if (active_out) std::memcpy(active_out, g->g.active.data(), g->g.active.size());
    return LSG_OK;
}

int lsg_grid_device_order(const lsg_grid* g, int64_t* dev_of_slot) {
    if (!g || !dev_of_slot) return fail(LSG_ERR_ARG, "null argument");
    std::memcpy(dev_of_slot, g->g.devOfSlot.data(), g->g.devOfSlot.size() * sizeof(int64_t));
    return LSG_OK;
}

int lsg_grid_stats(const lsg_grid* g, int64_t* out, int nout) {
    if (!g || !out) return fail(LSG_ERR_ARG, "null argument");
    planStats(g->g, g->g.plan, out, nout);
    return LSG_OK;
}

int lsg_op_stats(const lsg_op* op, int64_t* out, int nout) {
    if (!op || !out) return fail(LSG_ERR_ARG, "null argument");
    planStats(*op->G, *op->P, out, nout);
    return LSG_OK;
}

int lsg_grid_cells(lsg_grid* g, int64_t* ncells, double* centres, int32_t* depth_index) {
    if (!g || !ncells) return fail(LSG_ERR_ARG, "null argument");
    const Grid* cg = nullptr;
    if (int rc = cellGrid(g, cg)) return rc;
    std::vector<double> x;
    std::vector<int32_t> t;
    cellCentres(*cg, x, &t);
    *ncells = (int64_t)t.size();
    if (centres) std::memcpy(centres, x.data(), x.size() * sizeof(double));
    if (depth_index) std::memcpy(depth_index, t.data(), t.size() * sizeof(int32_t));
    return LSG_OK;
}

int lsg_grid_cell_boxes(lsg_grid* g, int32_t* level, int32_t* pos) {
    if (!g) return fail(LSG_ERR_ARG, "null argument");
    const Grid* cg = nullptr;
    if (int rc = cellGrid(g, cg)) return rc;
    const SparseTables& S = cg->sp;
    const int d = cg->d;
    const int nd = (int)cg->depths.size();
    for (int t = 0; t < nd; ++t)
        for (int64_t k = S.cellOff[t]; k < S.cellOff[t + 1]; ++k)
            for (int i = 0; i < d; ++i) {
                if (level) level[k * d + i] = cg->depths[t][i];
                if (pos) pos[k * d + i] = S.cellLo[(size_t)(k * d + i)];
            }
    return LSG_OK;
}

int64_t lsg_cell_entries(int d) { return (d < 1 || d > MAXD) ? -1 : cellEntries(d); }

int lsg_grid_destroy(lsg_grid* g) {
    delete g;
    return LSG_OK;
}

int lsg_op_create(lsg_grid* grid, const lsg_coef* coef, int precision, int batch, lsg_op** out) {
    if (!grid || !coef || !out || batch < 1) return fail(LSG_ERR_ARG, "bad arguments");
    if (precision != 64 && precision != 32) return fail(LSG_ERR_ARG, "precision must be 64 or 32");
    std::unique_ptr<lsg_op> op(new lsg_op());
    op->grid = grid;
    op->prec = precision;
    op->vs = precision == 64 ? 8 : 4;
    op->batch = batch;
    op->kind = coef->kind;
    op->cellKind = cellCoefKind(coef->kind);
    op->matKind = matCoefKind(coef->kind);
    op->G = &grid->g;
    if (op->cellKind) {
        int rc = cellGrid(grid, op->G);
        if (rc) return rc;
    }
    const Grid& g = *op->G;
    op->P = &g.plan;
    if (!g.sparse && batch <= singlePlanMaxBatch()) {
        // small batches: every prolongation dimension materialised, single-item apply
        // launches (more workgroups per wavefront level, a shorter critical path)
        std::string err;
        if (int rc = buildSinglePlan(grid->g, err)) return fail(rc, err);
        op->P = grid->g.planSingle.get();
    }
    const Plan& P = *op->P;
    if (coefTerms(*coef, g.d, op->nA, op->nC)) return fail(LSG_ERR_ARG, "bad coefficient");
    if (op->nA + op->nC == 0) return fail(LSG_ERR_ARG, "zero operator");
    op->general = !(op->nA == 1 && op->nC == 0);
    op->tabLen = (int64_t(1) << (g.maxLevel + 1)) - 2;
    op->sparse = g.sparse;
    if (!op->sparse && P.maxBox >= 65536)
        return fail(LSG_ERR_UNSUPPORTED, "depth boxes above 65535 points are not supported");
    // size classes of the apply launches (largest box of a family): a launch sizes its
    // LDS and points per thread for its largest box, so narrow classes keep the
    // occupancy of the smaller families up; small batches are launch-latency bound and
    // take fewer, wider classes
    const int64_t fine[12] = {128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 4096, 8192, 65536};
    const int64_t coarse[5] = {128, 512, 2048, 8192, 65536};
    const int64_t one[1] = {65536};
    // before autotuning: fine classes for batches of 16 or more, one class up to
    // oneClassMaxBatch(), else the coarse ones
    op->cs = batch >= 16 ? 0 : (batch <= oneClassMaxBatch() ? 2 : 1);
    for (int set = 0; set < 3; ++set)
    for (size_t L = 0; !op->sparse && L < P.apply.size(); ++L) {
        const int64_t* bounds = set == 0 ? fine : (set == 1 ? coarse : one);
        const int ncls = set == 0 ? 12 : (set == 1 ? 5 : 1);
        auto& S = op->sets[set];
        std::vector<std::vector<FamilyItem>> bucket(ncls);
        std::vector<int64_t> bmax(ncls, 0), wmax(ncls, 0);
        for (const FamilyItem& f : P.apply[L]) {
            int64_t mb = 0, mw = 0;
            for (int q = 0; q < f.nt; ++q) {
                const ApplyItem& a = P.applyItems[f.first + q];
                mb = std::max(mb, boxCount(a.T, g.d));
                int64_t w = 0;
                for (int i = 0; i < g.d; ++i) w += int64_t(1) << a.T[i];
                mw = std::max(mw, w);
            }
            if (!op->general) {
                // the separable apply kernel keeps the weights of the whole family:
                // along fd the levels T_fd .. T_fd + nt - 1 back to back
                const ApplyItem& a0 = P.applyItems[f.first];
                int64_t w = 0;
                for (int i = 0; i < g.d; ++i)
                    w += (i == f.fd) ? (int64_t(1) << (a0.T[i] + f.nt)) - (int64_t(1) << a0.T[i])
                                     : int64_t(1) << a0.T[i];
                mw = std::max(mw, w);
            }
            int c = 0;
            while (c < ncls - 1 && mb > bounds[c]) ++c;
            bucket[c].push_back(f);
            bmax[c] = std::max(bmax[c], mb);
            wmax[c] = std::max(wmax[c], mw);
        }
        for (int c = 0; c < ncls; ++c) {
            if (bucket[c].empty()) continue;
            // longest families first (work = points of all boxes of the family)
            std::stable_sort(bucket[c].begin(), bucket[c].end(), [&](const FamilyItem& x, const FamilyItem& y) {
                auto work = [&](const FamilyItem& f) {
                    int64_t w = 0;
                    for (int q = 0; q < f.nt; ++q) w += boxCount(P.applyItems[f.first + q].T, g.d);
                    return w;
                };
                return work(x) > work(y);
            });
            lsg_op::ApplyLaunch la;
            la.level = (int)L;
            la.cls = c;
            la.count = (int)bucket[c].size();
            la.maxBox = bmax[c];
            la.maxW = wmax[c];
            // instances per workgroup: about 4096 points of work per sweep, at most
            // 64 KB of LDS, exact fast division (G * box < 2^16)
            {
                const int64_t perInst = ((op->general ? 4 : 2) * la.maxBox + la.maxW) * (int64_t)op->vs;
                int64_t G = std::max<int64_t>(1, 1024 / la.maxBox);
                G = std::min<int64_t>(G, std::max<int64_t>(1, (48 * 1024) / perInst));
                G = std::min<int64_t>(G, 65535 / la.maxBox);
                G = std::min<int64_t>(G, MAXG);
                G = std::min<int64_t>(G, batch);
                la.G = (int)std::max<int64_t>(G, 1);
            }
            la.fam.reset(new DevMem());
            S.launches.push_back(std::move(la));
            S.nClasses = std::max(S.nClasses, c + 1);
            auto& dm = *S.launches.back().fam;
            HIPCHK(dm.alloc(bucket[c].size() * sizeof(FamilyItem)));
            HIPCHK(hipMemcpy(dm.p, bucket[c].data(), bucket[c].size() * sizeof(FamilyItem), hipMemcpyHostToDevice));
        }
    }
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    op->device = dev;
    int maxLds = 0;
    HIPCHK(hipDeviceGetAttribute(&maxLds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    op->maxLds = maxLds;
    // a class set with a launch beyond the LDS of a workgroup is dropped (the one-class set
    // combines the largest box and the largest weight table of a level)
    int64_t ldsNeed = 0;
    bool anySet = false;
    for (auto& S : op->sets) {
        bool fits = true;
        for (auto& L : S.launches) {
            const int64_t lds = applyLdsBytes(op.get(), L.maxBox, L.maxW, 512);
            ldsNeed = std::max(ldsNeed, lds);
            if (lds > maxLds) fits = false;
        }
        if (!fits) {
            S.launches.clear();
            S.nClasses = 0;
        }
        anySet |= S.nClasses > 0;
    }
    if (!op->sparse && !anySet)
        return fail(LSG_ERR_UNSUPPORTED, "largest depth box does not fit in LDS (" + std::to_string(ldsNeed) + " > " +
                                             std::to_string(maxLds) + " bytes)");
    if (!op->sparse && op->act().nClasses == 0)
        for (int set = 0; set < 3; ++set)
            if (op->sets[set].nClasses > 0) { op->cs = set; break; }
    if (!op->sparse && (int64_t)(2 * P.maxBox * op->vs) > maxLds)
        return fail(LSG_ERR_UNSUPPORTED, "largest depth box does not fit in LDS");
    HIPCHK(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
    for (auto& e : op->ev) HIPCHK(hipEventCreate(&e));
    // items
    HIPCHK(op->dNodalItems.alloc(P.nodal.size() * sizeof(NodalItem)));
    HIPCHK(hipMemcpy(op->dNodalItems.p, P.nodal.data(), P.nodal.size() * sizeof(NodalItem), hipMemcpyHostToDevice));
    for (const auto& v0 : P.prol) {
        // longest chains first; the largest box of the sweep sizes the LDS
        std::vector<ProlItem> v(v0);
        auto boxes = [&](const ProlItem& it, bool largest) {
            int64_t w = 0, m = 0;
            int32_t T[MAXD];
            std::memcpy(T, it.T, sizeof(T));
            for (int q = 0; q < it.nt; ++q) {
                T[it.dd] = it.T[it.dd] + q;
                const int64_t c = boxCount(T, g.d);
                w += c;
                m = std::max(m, c);
            }
            return largest ? m : w;
        };
        std::stable_sort(v.begin(), v.end(),
                         [&](const ProlItem& x, const ProlItem& y) { return boxes(x, false) > boxes(y, false); });
        int64_t mb = 1, mc = 1;
        for (const ProlItem& it : v) {
            mb = std::max(mb, boxes(it, true));
            mc = std::max(mc, boxes(it, false));
        }
        op->prolMax.push_back(mb);
        op->prolChain.push_back(mc);
        op->dProl.emplace_back(new DevMem());
        HIPCHK(op->dProl.back()->alloc(v.size() * sizeof(ProlItem)));
        if (!v.empty()) HIPCHK(hipMemcpy(op->dProl.back()->p, v.data(), v.size() * sizeof(ProlItem), hipMemcpyHostToDevice));
    }
    HIPCHK(op->dShapes.alloc(std::max<size_t>(1, P.shapes.size()) * sizeof(BoxShape)));
    if (!P.shapes.empty())
        HIPCHK(hipMemcpy(op->dShapes.p, P.shapes.data(), P.shapes.size() * sizeof(BoxShape), hipMemcpyHostToDevice));
    HIPCHK(op->dApplyItems.alloc(P.applyItems.size() * sizeof(ApplyItem)));
    HIPCHK(hipMemcpy(op->dApplyItems.p, P.applyItems.data(), P.applyItems.size() * sizeof(ApplyItem),
                     hipMemcpyHostToDevice));

    HIPCHK(op->dSumItems.alloc(P.sum.size() * sizeof(SumItem)));
    HIPCHK(hipMemcpy(op->dSumItems.p, P.sum.data(), P.sum.size() * sizeof(SumItem), hipMemcpyHostToDevice));
    HIPCHK(op->dSumList.alloc(P.sumList.size() * sizeof(int64_t)));
    if (!P.sumList.empty())
        HIPCHK(hipMemcpy(op->dSumList.p, P.sumList.data(), P.sumList.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    // buffers
    const size_t vs = op->vs;
    HIPCHK(op->nodal.alloc((size_t)g.nml * batch * vs));
    HIPCHK(op->cbuf.alloc((size_t)P.caseSize * batch * vs));
    HIPCHK(op->wbuf.alloc((size_t)P.wSize * batch * vs));
    HIPCHK(op->part.alloc((size_t)P.partSize * batch * vs));
    HIPCHK(op->xs.alloc((size_t)g.n * batch * vs));
    HIPCHK(op->ys.alloc((size_t)g.n * batch * vs));
    HIPCHK(hipMemset(op->xs.p, 0, op->xs.bytes));
    HIPCHK(op->coefTab.alloc((size_t)(op->nA + op->nC) * g.d * op->tabLen * batch * vs));
    if (op->sparse) {
        // gather tables of the sparse path, single-item families per wavefront level, scratch
        const SparseTables& T = g.sp;
        auto up = [&](const void* data, size_t bytes) -> const void* {
            op->spTables.emplace_back(new DevMem());
            DevMem& m = *op->spTables.back();
            if (m.alloc(bytes) != hipSuccess) return nullptr;
            if (bytes && hipMemcpy(m.p, data, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
            return m.p;
        };
#define LSG_UP(field, vec)                                                                    \
    if (!(op->sd.field = (decltype(op->sd.field))up(vec.data(), vec.size() * sizeof(vec[0])))) \
        return fail(LSG_ERR_DEVICE, "upload of the sparse tables failed");
        LSG_UP(nodPtr, T.nodPtr) LSG_UP(nodCol, T.nodCol) LSG_UP(nodVal, T.nodVal)
        LSG_UP(prolOff, T.prolOff) LSG_UP(prolAB, T.prolAB)
        LSG_UP(cellOff, T.cellOff) LSG_UP(cellCorner, T.cellCorner) LSG_UP(cellLo, T.cellLo)
        LSG_UP(cellColOff, T.cellColOff) LSG_UP(cellCol, T.cellCol)
        LSG_UP(rcPtr, T.rcPtr) LSG_UP(rcCell, T.rcCell) LSG_UP(rcCorner, T.rcCorner) LSG_UP(rcY, T.rcY)
        LSG_UP(restOff, T.restOff) LSG_UP(restSLR, T.restSLR)
        LSG_UP(convPtr, T.convPtr) LSG_UP(convCol, T.convCol) LSG_UP(convVal, T.convVal)
        {
            const std::vector<float> cf(T.convVal.begin(), T.convVal.end()), nf(T.nodVal.begin(), T.nodVal.end());
            LSG_UP(convValF, cf) LSG_UP(nodValF, nf)
        }
        {
            // 16-bit index tables when every box and every depth's cell list fits them
            int64_t big = 0;
            for (size_t t = 0; t < g.depths.size(); ++t)
                big = std::max({big, g.boxSize[t], T.cellOff[t + 1] - T.cellOff[t]});
            op->sd.cellCorner16 = op->sd.cellCol16 = op->sd.restSLR16 = op->sd.convCol16 = nullptr;
            if (big < 32768 && !envInt("LSG_SP_IX32", 0)) {
                auto narrow = [](const std::vector<int32_t>& v) { return std::vector<int16_t>(v.begin(), v.end()); };
                const std::vector<int16_t> cc16 = narrow(T.cellCorner), cl16 = narrow(T.cellCol),
                                           rs16 = narrow(T.restSLR), cv16 = narrow(T.convCol);
                LSG_UP(cellCorner16, cc16) LSG_UP(cellCol16, cl16) LSG_UP(restSLR16, rs16) LSG_UP(convCol16, cv16)
            }
        }
        LSG_UP(pos, T.pos) LSG_UP(tOfDev, T.tOfDev) LSG_UP(depthT, T.depthT)
        LSG_UP(boxOff, g.boxOff) LSG_UP(boxSize, g.boxSize) LSG_UP(blkOff, g.blkOff) LSG_UP(blkSize, g.blkSize)
#undef LSG_UP
        op->sd.d = g.d;
        op->sd.nd = (int)g.depths.size();
        // the cell-wise apply keeps an item's point values in LDS: the box and, after the
        // first restriction, the coarser box; items are launched by the size of that scratch
        // (SP_SMALL up to 8 KB per instance, four instances per workgroup; SP_MEDIUM up to
        // 48 KB; SP_LARGE up to the device's LDS; larger ones in zbuf)
        const int64_t ldsMax = std::min<int64_t>(op->maxLds, envInt("LSG_SP_LDS_BYTES", (int)op->maxLds));
        const int64_t smallMax = envInt("LSG_SP_SMALL_BYTES", 8192), mediumMax = envInt("LSG_SP_MEDIUM_BYTES", 49152);
        bool anyGlobal = false;
        for (const auto& lev : P.apply) {
            std::vector<FamilyItem> byCls[SP_NCLASS];
            int64_t stride[SP_NCLASS] = {};
            for (const FamilyItem& f : lev) {
                const ApplyItem& it = P.applyItems[(size_t)f.first];
                const int64_t need = g.boxSize[it.tidx] + (it.k >= 1 ? g.boxSize[it.ctidx[1]] : 0);
                const int64_t bytes = need * (int64_t)vs;
                const int c = bytes <= std::min(smallMax, ldsMax / 4) ? SP_SMALL
                              : bytes <= std::min(mediumMax, ldsMax)     ? SP_MEDIUM
                              : bytes <= ldsMax                          ? SP_LARGE
                                                                         : SP_GLOBAL;
                byCls[c].push_back(f);
                stride[c] = std::max(stride[c], need);
            }
            for (int c = 0; c < SP_NCLASS; ++c) {
                if (byCls[c].empty()) continue;
                std::unique_ptr<DevMem> m(new DevMem());
                HIPCHK(m->alloc(byCls[c].size() * sizeof(FamilyItem)));
                HIPCHK(hipMemcpy(m->p, byCls[c].data(), byCls[c].size() * sizeof(FamilyItem), hipMemcpyHostToDevice));
                op->spLevels.push_back({(int)byCls[c].size(), c, c == SP_GLOBAL ? 0 : stride[c], std::move(m)});
                anyGlobal = anyGlobal || c == SP_GLOBAL;
            }
        }
        // the per-item scratch in HBM serves the stored element matrices, d > SP_CELL_MAXD and
        // items whose point values exceed the LDS; otherwise the apply keeps it all in LDS
        const bool needZ = op->matKind || g.d > SP_CELL_MAXD || anyGlobal;
        HIPCHK(op->zbuf.alloc(needZ ? (size_t)P.zSize * batch * vs : vs));
        op->nCells = T.cellOff.back();
        HIPCHK(op->cellCoef.alloc((size_t)std::max<int64_t>(1, 2 * op->nCells) * batch * vs));
        if (op->matKind) {
            op->nEnt = cellEntries(g.d);
            const size_t bytes = (size_t)std::max<int64_t>(1, op->nCells * op->nEnt) * batch * vs;
            size_t freeB = 0, totalB = 0;
            HIPCHK(hipMemGetInfo(&freeB, &totalB));
            if (bytes > freeB)
                return fail(LSG_ERR_CAPACITY, "stored element matrices need " + std::to_string(bytes) +
                                                  " bytes of device memory, " + std::to_string(freeB) + " are free");
            HIPCHK(op->cellMat.alloc(bytes));
        }
        std::vector<uint8_t> act(g.n);
        for (int64_t s0 = 0; s0 < g.n; ++s0) act[g.devOfSlot[s0]] = g.active[s0];
        HIPCHK(op->activeDev.alloc((size_t)g.n));
        HIPCHK(hipMemcpy(op->activeDev.p, act.data(), (size_t)g.n, hipMemcpyHostToDevice));
    }
    *out = op.release();
    for (int b = 0; b < batch; ++b) {
        int rc = lsg_op_set_coef(*out, b, coef);
        if (rc) { lsg_op_destroy(*out); *out = nullptr; return rc; }
    }
    return LSG_OK;
}

int lsg_op_set_coef(lsg_op* op, int instance, const lsg_coef* coef) {
    if (!op || !coef || instance < 0 || instance >= op->batch) return fail(LSG_ERR_ARG, "bad arguments");
    int nA, nC;
    const Grid& g = *op->G;
    if (coefTerms(*coef, g.d, nA, nC) || nA != op->nA || nC != op->nC || cellCoefKind(coef->kind) != op->cellKind ||
        matCoefKind(coef->kind) != op->matKind)
        return fail(LSG_ERR_ARG, "coefficient family differs from the operator's");
    if (op->matKind) {
        // stored element matrices, straight into cellMat
        std::vector<double> E;
        if (int rc = cellMatrices(g, *coef, E)) return rc;
        const size_t per = (size_t)(op->nCells * op->nEnt);
        if (E.size() != per) return fail(LSG_ERR_ARG, "cell count mismatch");
        if (op->prec == 64) {
            HIPCHK(hipMemcpy(op->cellMat.as<double>() + per * instance, E.data(), per * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> ef(E.begin(), E.end());
            HIPCHK(hipMemcpy(op->cellMat.as<float>() + per * instance, ef.data(), per * 4, hipMemcpyHostToDevice));
        }
        op->diagValid = false;
        op->cellCoefValid = true;
        return LSG_OK;
    }
    if (op->cellKind) {
        // general coefficient: (a, c) per cell sampled on the host, straight into cellCoef
        std::vector<double> ac;
        if (int rc = cellValues(g, *coef, ac)) return rc;
        const size_t per = (size_t)2 * op->nCells;
        if (ac.size() != per) return fail(LSG_ERR_ARG, "cell count mismatch");
        {   // device layout: the a values of all cells, then the c values
            std::vector<double> planar(per);
            for (int64_t c = 0; c < op->nCells; ++c) {
                planar[(size_t)c] = ac[(size_t)2 * c];
                planar[(size_t)(op->nCells + c)] = ac[(size_t)2 * c + 1];
            }
            ac.swap(planar);
        }
        if (op->prec == 64) {
            HIPCHK(hipMemcpy(op->cellCoef.as<double>() + per * instance, ac.data(), per * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> af(ac.begin(), ac.end());
            HIPCHK(hipMemcpy(op->cellCoef.as<float>() + per * instance, af.data(), per * 4, hipMemcpyHostToDevice));
        }
        op->diagValid = false;
        op->cellCoefValid = true;
        return LSG_OK;
    }
    std::vector<double> tab;
    coefTables(*coef, g.d, g.maxLevel, op->tabLen, tab);
    const size_t per = tab.size();
    if (op->prec == 64) {
        HIPCHK(hipMemcpy(op->coefTab.as<double>() + per * instance, tab.data(), per * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tf(tab.begin(), tab.end());
        HIPCHK(hipMemcpy(op->coefTab.as<float>() + per * instance, tf.data(), per * 4, hipMemcpyHostToDevice));
    }
    op->diagValid = false;
    op->cellCoefValid = false;
    return LSG_OK;
}

// Coefficients of instances first .. first + count - 1 in one call: the coefficient tables are
// computed on up to 16 host threads and uploaded with one copy (general coefficients, which
// are sampled per cell, go through lsg_op_set_coef one by one).
int lsg_op_set_coefs(lsg_op* op, int first, int count, const lsg_coef* coefs) {
    if (!op || !coefs || first < 0 || count < 0 || first + count > op->batch) return fail(LSG_ERR_ARG, "bad arguments");
    if (count == 0) return LSG_OK;
    const Grid& g = *op->G;
    if (op->cellKind) {
        for (int b = 0; b < count; ++b)
            if (int rc = lsg_op_set_coef(op, first + b, &coefs[b])) return rc;
        return LSG_OK;
    }
    for (int b = 0; b < count; ++b) {
        int nA, nC;
        if (coefTerms(coefs[b], g.d, nA, nC) || nA != op->nA || nC != op->nC || cellCoefKind(coefs[b].kind))
            return fail(LSG_ERR_ARG, "coefficient family differs from the operator's");
    }
    const size_t per = (size_t)(op->nA + op->nC) * g.d * op->tabLen;
    std::vector<double> tab(per * count);
    const int nth = std::max(1, std::min<int>({16, count, (int)std::thread::hardware_concurrency()}));
    auto work = [&](int w) {
        std::vector<double> one;
        for (int b = w; b < count; b += nth) {
            coefTables(coefs[b], g.d, g.maxLevel, op->tabLen, one);
            std::copy(one.begin(), one.end(), tab.begin() + per * b);
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < nth; ++w) th.emplace_back(work, w);
    work(0);
    for (auto& t : th) t.join();
    if (op->prec == 64) {
        HIPCHK(hipMemcpy(op->coefTab.as<double>() + per * first, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tf(tab.begin(), tab.end());
        HIPCHK(hipMemcpy(op->coefTab.as<float>() + per * first, tf.data(), tf.size() * 4, hipMemcpyHostToDevice));
    }
    op->diagValid = false;
    op->cellCoefValid = false;
    return LSG_OK;
}

int lsg_op_batch(const lsg_op* op) { return op ? op->batch : -1; }

int lsg_op_set_precond(lsg_op* op, int kind) {
    if (!op) return fail(LSG_ERR_ARG, "null operator");
    if (kind != LSG_PRECOND_DIAG && kind != LSG_PRECOND_DEPTH_BLOCK) return fail(LSG_ERR_ARG, "unknown preconditioner");
    if (kind == LSG_PRECOND_DEPTH_BLOCK) {
        if (int rc = ensureBlockPre(op)) return rc;
    }
    op->precKind = kind;
    return LSG_OK;
}

int lsg_op_set_block(lsg_op* op, int stage, int threads) {
    if (!op || stage != 0) return fail(LSG_ERR_ARG, "only stage 0 (apply) is tunable");
    if (threads < 0 || threads > 512 || threads % 64 != 0)
        return fail(LSG_ERR_ARG, "threads must be a multiple of 64 up to 512");
    return chooseThreads(op, threads);
}

int lsg_matvec(lsg_op* op, const void* x, void* y, int layout) {
    if (!op || !x || !y) return fail(LSG_ERR_ARG, "null argument");
    int rc = ensureThreads(op);
    if (rc) return rc;
    if (layout == LSG_DEVICE) {
        rc = matvecDev(op, x, y, nullptr);
        if (rc) return rc;
        HIPCHK(hipStreamSynchronize(op->stream));
        return LSG_OK;
    }
    if ((rc = upload(op, (const double*)x, op->xs.p))) return rc;
    if ((rc = matvecDev(op, op->xs.p, op->ys.p, nullptr))) return rc;
    return download(op, op->ys.p, (double*)y);
}

int lsg_precond_diag(lsg_op* op, void* z, int layout) {
    if (!op || !z) return fail(LSG_ERR_ARG, "null argument");
    int rc = ensureSolver(op);
    if (rc) return rc;
    if ((rc = computeDiag(op))) return rc;
    if (layout == LSG_DEVICE) {
        HIPCHK(hipMemcpyAsync(z, op->D.p, (size_t)op->n() * op->batch * op->vs, hipMemcpyDeviceToDevice, op->stream));
        HIPCHK(hipStreamSynchronize(op->stream));
        return LSG_OK;
    }
    return download(op, op->D.p, (double*)z);
}

int lsg_cg_solve(lsg_op* op, const void* f, void* x, double eps, int stop_rule, int maxit, lsg_solve_info* info,
                 int layout) {
    if (!op || !f || !x || maxit < 0) return fail(LSG_ERR_ARG, "bad arguments");
    int rc = ensureSolver(op);
    if (rc) return rc;
    const int64_t tot = op->n() * op->batch;
    const void* fD = f;
    void* xD = x;
    if (layout == LSG_HOST_SLOT) {
        // f in op->f (V), x in op->x (V)
        if ((rc = upload(op, (const double*)f, op->f.p))) return rc;
        if ((rc = upload(op, (const double*)x, op->x.p))) return rc;
        fD = op->f.p;
        xD = op->x.p;
    }
    std::vector<CGState> st;
    const bool precond = !(stop_rule & LSG_CG_NO_PRECOND);
    const int rule = stop_rule & ~LSG_CG_NO_PRECOND;
    if (rule != LSG_STOP_REF_ABS && rule != LSG_STOP_REL) return fail(LSG_ERR_ARG, "unknown stopping rule");
    if (op->prec == 64)
        rc = cgDevice<double>(op, (const double*)fD, xD, true, eps, rule, maxit, st, false, nullptr, nullptr, precond);
    else
        rc = cgDevice<float>(op, (const float*)fD, xD, false, eps, rule, maxit, st, false, nullptr, nullptr, precond);
    if (rc) return rc;
    (void)tot;
    if (layout == LSG_HOST_SLOT && (rc = download(op, op->x.p, (double*)x))) return rc;
    bool nan = false, noconv = false;
    for (int b = 0; b < op->batch; ++b) {
        if (info) {
            info[b].iterations = st[b].it;
            info[b].outer = 0;
            info[b].final_delta = st[b].delta;
            info[b].rel_residual = st[b].ff > 0 ? std::sqrt(st[b].rr / st[b].ff) : std::sqrt(st[b].rr);
            info[b].converged = st[b].converged;
        }
        nan |= st[b].converged < 0;
        noconv |= st[b].converged == 0;
    }
    if (nan) return fail(LSG_ERR_NAN, "NaN in CG");
    if (noconv) return fail(LSG_ERR_NOCONV, "CG stopped at maxit before the stopping rule was met");
    return LSG_OK;
}

int lsg_cg_solve_noalloc(lsg_op* op, const void* f, void* x, double eps, int stop_rule, int maxit,
                         lsg_solve_info* info) {
    if (!op || !f || !x || maxit < 0) return fail(LSG_ERR_ARG, "bad arguments");
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != op->device) HIPCHK(hipSetDevice(op->device));
    std::vector<CGState> st;
    int rc = op->prec == 64 ? cgNoalloc<double>(op, f, x, eps, stop_rule, maxit, st)
                            : cgNoalloc<float>(op, f, x, eps, stop_rule, maxit, st);
    if (cur != op->device) (void)hipSetDevice(cur);
    if (rc) return rc;
    bool nan = false, noconv = false;
    for (int b = 0; b < op->batch; ++b) {
        if (info) {
            info[b].iterations = st[b].it;
            info[b].outer = 0;
            info[b].final_delta = st[b].delta;
            info[b].rel_residual = st[b].ff > 0 ? std::sqrt(st[b].rr / st[b].ff) : std::sqrt(st[b].rr);
            info[b].converged = st[b].converged;
        }
        nan |= st[b].converged < 0;
        noconv |= st[b].converged == 0;
    }
    if (nan) return fail(LSG_ERR_NAN, "NaN in CG");
    if (noconv) return fail(LSG_ERR_NOCONV, "CG stopped at maxit before the stopping rule was met");
    return LSG_OK;
}

int lsg_ir_solve(lsg_op* op64, lsg_op* op32, const void* f, void* x, double rel_tol, double inner_rel_tol, int maxit,
                 lsg_solve_info* info, int layout) {
    if (!op64 || !op32 || !f || !x) return fail(LSG_ERR_ARG, "null argument");
    if (op64->prec != 64 || op32->prec != 32 || op64->G != op32->G || op64->batch != op32->batch)
        return fail(LSG_ERR_ARG, "need an FP64 and an FP32 operator on the same grid and batch");
    int rc;
    if ((rc = ensureSolver(op64))) return rc;
    std::vector<IrGroup> G(1);
    G[0].op64 = op64;
    G[0].op32 = op32;
    if (layout == LSG_HOST_SLOT) {
        if ((rc = upload(op64, (const double*)f, op64->f.p)) || (rc = upload(op64, (const double*)x, op64->x.p)))
            return rc;
        G[0].f = op64->f.as<double>();
        G[0].x = op64->x.as<double>();
    } else {
        G[0].f = (double*)f;
        G[0].x = (double*)x;
    }
    rc = irSolveGroups(G, rel_tol, inner_rel_tol, maxit);
    if (rc) return rc;
    if (layout == LSG_HOST_SLOT && (rc = download(op64, op64->x.p, (double*)x))) return rc;
    bool noconv = false;
    for (int b = 0; b < op64->batch; ++b) {
        if (info) {
            info[b].iterations = G[0].iters[b];
            info[b].outer = G[0].outer[b];
            info[b].final_delta = 0.0;
            info[b].rel_residual = G[0].rel[b];
            info[b].converged = G[0].conv[b];
        }
        noconv |= !G[0].conv[b];
    }
    if (noconv) return fail(LSG_ERR_NOCONV, "iterative refinement stopped at maxit sweeps before rel_tol was met");
    return LSG_OK;
}

int lsg_ir_solve_many(int ngroups, lsg_op* const* op64, lsg_op* const* op32, const void* const* f, void* const* x,
                      double rel_tol, double inner_rel_tol, int maxit, lsg_solve_info* const* info) {
    if (ngroups < 1 || !op64 || !op32 || !f || !x) return fail(LSG_ERR_ARG, "bad arguments");
    std::vector<IrGroup> G(ngroups);
    for (int q = 0; q < ngroups; ++q) {
        lsg_op* a = op64[q];
        lsg_op* b = op32[q];
        if (!a || !b || !f[q] || !x[q]) return fail(LSG_ERR_ARG, "null operator or vector in group " + std::to_string(q));
        if (a->prec != 64 || b->prec != 32 || a->G != b->G || a->batch != b->batch)
            return fail(LSG_ERR_ARG, "group " + std::to_string(q) +
                                         ": need an FP64 and an FP32 operator on the same grid and batch");
        for (int p = 0; p < q; ++p)
            if (op64[p] == a || op32[p] == b || op64[p] == b || op32[p] == a)
                return fail(LSG_ERR_ARG, "an operator appears in two groups");
        G[q].op64 = a;
        G[q].op32 = b;
        G[q].f = (double*)f[q];
        G[q].x = (double*)x[q];
    }
    int rc = irSolveGroups(G, rel_tol, inner_rel_tol, maxit);
    if (rc) return rc;
    bool noconv = false;
    for (int q = 0; q < ngroups; ++q)
        for (int b = 0; b < G[q].B; ++b) {
            if (info && info[q]) {
                lsg_solve_info& o = info[q][b];
                o.iterations = G[q].iters[b];
                o.outer = G[q].outer[b];
                o.final_delta = 0.0;
                o.rel_residual = G[q].rel[b];
                o.converged = G[q].conv[b];
            }
            noconv |= !G[q].conv[b];
        }
    if (noconv) return fail(LSG_ERR_NOCONV, "iterative refinement stopped at maxit sweeps before rel_tol was met");
    return LSG_OK;
}

int lsg_rhs_sines(lsg_op* op, int nterms, const double* eta, const int32_t* m, void* b, int layout) {
    if (!op || nterms < 1 || !eta || !m || !b) return fail(LSG_ERR_ARG, "bad arguments");
    const Grid& g = *op->G;
    DevMem& dEta = op->rhsEta;
    DevMem& dM = op->rhsM;
    if (dEta.bytes < (size_t)nterms * op->batch * 8) HIPCHK(dEta.alloc((size_t)nterms * op->batch * 8));
    if (dM.bytes < (size_t)nterms * 4) HIPCHK(dM.alloc((size_t)nterms * 4));
    HIPCHK(hipStreamSynchronize(op->stream));   // an earlier call's kernels may still read them
    HIPCHK(hipMemcpy(dEta.p, eta, (size_t)nterms * op->batch * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dM.p, m, (size_t)nterms * 4, hipMemcpyHostToDevice));
    void* dst = layout == LSG_DEVICE ? b : op->ys.p;
    const int32_t* pos = op->sparse ? op->sd.pos : nullptr;
    const uint8_t* act = op->sparse ? op->activeDev.as<uint8_t>() : nullptr;
    if (op->prec == 64)
        launch_rhs_sines<double>(op->dSumItems.as<SumItem>(), (int)op->P->sum.size(), g.d, nterms, dEta.as<double>(),
                                 dM.as<int32_t>(), (double*)dst, g.n, op->batch, op->stream, pos, act);
    else
        launch_rhs_sines<float>(op->dSumItems.as<SumItem>(), (int)op->P->sum.size(), g.d, nterms, dEta.as<double>(),
                                dM.as<int32_t>(), (float*)dst, g.n, op->batch, op->stream, pos, act);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(op->stream));
    if (layout == LSG_DEVICE) return LSG_OK;
    return download(op, op->ys.p, (double*)b);
}

int lsg_rhs_nodes(const lsg_op* op, int64_t* count, double* nodes) {
    if (!op || !count) return fail(LSG_ERR_ARG, "null argument");
    std::vector<double> x;
    rhsNodes(op->G->maxLevel, x);
    *count = (int64_t)x.size();
    if (nodes) std::memcpy(nodes, x.data(), x.size() * sizeof(double));
    return LSG_OK;
}

int lsg_rhs_separable(lsg_op* op, int nterms, const double* vals, const double* w, void* b, int layout) {
    if (!op || nterms < 1 || !vals || !w || !b) return fail(LSG_ERR_ARG, "bad arguments");
    const Grid& g = *op->G;
    const int d = g.d;
    const int L = g.maxLevel;
    const int64_t nq = (int64_t(1) << L) * 4;
    const int64_t tabLen = (int64_t(1) << L) - 1;
    std::vector<double> tab((size_t)nterms * d * tabLen);
    for (int64_t q = 0; q < (int64_t)nterms * d * nq; ++q)
        if (!std::isfinite(vals[q])) return fail(LSG_ERR_NAN, "non-finite factor value");
    for (int j = 0; j < nterms; ++j)
        for (int i = 0; i < d; ++i)
            factorMoments(L, vals + ((size_t)j * d + i) * nq, &tab[((size_t)j * d + i) * tabLen]);
    DevMem& dW = op->rhsEta;
    DevMem& dT = op->rhsM;
    HIPCHK(hipStreamSynchronize(op->stream));   // an earlier call's kernels may still read them
    if (dW.bytes < (size_t)nterms * op->batch * 8) HIPCHK(dW.alloc((size_t)nterms * op->batch * 8));
    if (dT.bytes < tab.size() * 8) HIPCHK(dT.alloc(tab.size() * 8));
    HIPCHK(hipMemcpy(dW.p, w, (size_t)nterms * op->batch * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dT.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
    void* dst = layout == LSG_DEVICE ? b : op->ys.p;
    const int32_t* pos = op->sparse ? op->sd.pos : nullptr;
    const uint8_t* act = op->sparse ? op->activeDev.as<uint8_t>() : nullptr;
    if (op->prec == 64)
        launch_rhs_tab<double>(op->dSumItems.as<SumItem>(), (int)op->P->sum.size(), d, nterms, dW.as<double>(),
                               dT.as<double>(), tabLen, (double*)dst, g.n, op->batch, op->stream, pos, act);
    else
        launch_rhs_tab<float>(op->dSumItems.as<SumItem>(), (int)op->P->sum.size(), d, nterms, dW.as<double>(),
                              dT.as<double>(), tabLen, (float*)dst, g.n, op->batch, op->stream, pos, act);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(op->stream));
    if (layout == LSG_DEVICE) return LSG_OK;
    return download(op, op->ys.p, (double*)b);
}

int lsg_hierarchize(lsg_op* op, const void* u, void* hier, int layout) { return transformAbi(op, 0, u, hier, layout); }

int lsg_dehierarchize(lsg_op* op, const void* hier, void* u, int layout) { return transformAbi(op, 1, hier, u, layout); }

int lsg_prew_by_nodal(lsg_op* op, const void* u, void* prew, int layout) { return transformAbi(op, 2, u, prew, layout); }

int lsg_nodal_by_prew(lsg_op* op, const void* prew, void* u, int layout) { return transformAbi(op, 3, prew, u, layout); }

int lsg_energy_norm(lsg_op* op, const void* u, double* out, int layout) {
    if (!op || !u || !out) return fail(LSG_ERR_ARG, "null argument");
    int rc = ensureTransform(op);
    if (rc) return rc;
    if ((rc = ensureThreads(op))) return rc;
    TrDev& t = *op->tr;
    const double* in = (const double*)u;
    if (layout != LSG_DEVICE) {
        if ((rc = upload(op, (const double*)u, t.work2.p))) return rc;
        in = t.work2.as<double>();
    }
    // v = prew(u) (H1_Norm.cpp:24 calcPrewByNodal), Av (:27), sqrt(v . Av) over the points (:29-37)
    if ((rc = transformDev(op, 2, in, t.work.as<double>()))) return rc;
    if ((rc = matvecDev(op, t.work.p, op->ys.p, nullptr))) return rc;
    launch_tr_energy(t.work.as<double>(), op->ys.as<double>(), op->n(), t.norms.as<double>(), op->batch, op->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, t.norms.p, (size_t)op->batch * 8, hipMemcpyDeviceToHost, op->stream));
    HIPCHK(hipStreamSynchronize(op->stream));
    return LSG_OK;
}

int lsg_vector_norms(lsg_op* op, const void* x, double* out, int layout) {
    if (!op || !x || !out) return fail(LSG_ERR_ARG, "null argument");
    const Grid& g = *op->G;
    const void* xv = x;
    if (layout != LSG_DEVICE) {
        int rc = upload(op, (const double*)x, op->xs.p);
        if (rc) return rc;
        xv = op->xs.p;
    }
    int64_t dofs = 0;
    for (uint8_t a : g.active) dofs += a ? 1 : 0;
    if (g.active.empty()) dofs = g.n;
    const uint8_t* act = op->sparse ? op->activeDev.as<uint8_t>() : nullptr;
    DevMem& res = op->norms2;
    if (!res.p) HIPCHK(res.alloc((size_t)op->batch * 2 * sizeof(double)));
    if (op->prec == 64)
        launch_vec_norms<double>((const double*)xv, g.n, act, (double)std::max<int64_t>(dofs, 1), res.as<double>(),
                                 op->batch, op->stream);
    else
        launch_vec_norms<float>((const float*)xv, g.n, act, (double)std::max<int64_t>(dofs, 1), res.as<double>(),
                                op->batch, op->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, res.p, (size_t)op->batch * 2 * sizeof(double), hipMemcpyDeviceToHost, op->stream));
    HIPCHK(hipStreamSynchronize(op->stream));
    return LSG_OK;
}

int lsg_last_timing(const lsg_op* op, double* out, int nout) {
    if (!op || !out || !op->timed) return fail(LSG_ERR_ARG, "no timed matvec");
    HIPCHK(hipEventSynchronize(op->ev[3]));
    float tot = 0.f, ap = 0.f;
    HIPCHK(hipEventElapsedTime(&tot, op->ev[0], op->ev[3]));
    HIPCHK(hipEventElapsedTime(&ap, op->ev[1], op->ev[2]));
    double v[3] = {tot, ap, (double)op->act().launches.size()};
    for (int i = 0; i < nout && i < 3; ++i) out[i] = v[i];
    return LSG_OK;
}

int lsg_op_destroy(lsg_op* op) {
    if (!op) return LSG_OK;
    if (op->stream) (void)hipStreamSynchronize(op->stream);
    for (auto& e : op->ev)
        if (e) (void)hipEventDestroy(e);
    if (op->chunkExec) (void)hipGraphExecDestroy(op->chunkExec);
    hipStream_t s = op->stream;
    delete op;
    if (s) (void)hipStreamDestroy(s);
    return LSG_OK;
}

}  // extern "C"
