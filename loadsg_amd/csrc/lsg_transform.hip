// lsg_transform.hip -- basis transformations between nodal values, hierarchical
// surpluses and prewavelet coefficients on gfx950, and the energy norms built on them
// (the reference's BasisTransformations/BasisTransformations.cpp and
// applications/H1_Norm.cpp).
//
//   * hierarchization / dehierarchization of a sparse-grid vector: point-parallel passes
//     over a host-built table of every point's poles along every dimension (the 1D
//     sweeps of CalcHierarchicalBasis / CalcNodalBasis, BasisTransformations.cpp:174-198,
//     390-409), one launch per dimension, and for the dehierarchization one per
//     (dimension, depth along it), coarse to fine, in place;
// The prewavelet steps work on the dense box G_T of one depth T (the level-T full grid,
// all points of depth <= T), staged in LDS by one workgroup per (depth, group of
// instances):
//   * prewavelets from nodal values (calcPrewByNodal, :486-545): depths are taken by
//     |T|_1 descending.  For one |T|_1 every depth box is dehierarchized from the
//     running hierarchical representation, solved line by line with the prewavelet
//     line matrix (PrewaveletMatrixHomogen), and its depth-T entries are the prewavelet
//     coefficients.  Their hierarchical representation on G_T (CalcHierbyPrew,
//     :321-353) goes to a scratch box and is subtracted point by point in a second
//     launch.  Depths with equal |T|_1 are incomparable, and the prewavelet part of
//     depth T of a function in the span of depths <= T' vanishes unless T <= T', so
//     the depths of one |T|_1 do not depend on each other;
//   * nodal values from prewavelets (calcNodalByPrew, :670-707): every depth's
//     hierarchical representation in one launch, a gather-sum per point, then
//     dehierarchization.
// On adaptive grids (hanging nodes, incomplete boxes) the box G_T keeps every level-T
// position with a status (not stored, hanging, active), and the reference's own order of
// work is kept:
//   * prewavelets from nodal values: one launch per depth, in the reference's order
//     (|T|_1 descending, ties in reverse first appearance).  The workgroup dehierarchizes
//     the running hierarchical representation on its box (only active entries change),
//     solves the lines of every connected component of the depth's points with the
//     factors of its cut line matrices (GetPrewaveletCoefficients,
//     BasisTransformations.cpp:100-172; only active entries are written), writes the
//     prewavelets, forms their hierarchical representation on the box (scatter onto
//     stored active entries, hierarchization of active entries) and subtracts it from
//     the running representation in place;
//   * nodal values from prewavelets: as on regular grids, with the same masks.
// Lines along one dimension are handled by one thread each (sequential recurrences).
// The boxes are small (at most a few thousand values), so a workgroup takes the boxes
// of one depth for several instances (up to 32 KB of LDS) to give its threads enough
// lines; a transform is two launches per |T|_1 level.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lsg_kernels.hpp"

namespace lsg {

namespace {

enum TrMode { TR_LEVEL = 2, TR_SYNTH = 3 };

// 1D hierarchization of a line of level t (n = 2^t - 1 values at stride s, 0 outside):
// finest level first, v_q -= (v_{q-h} + v_{q+h}) / 2 (CalcHierarchicalBasis,
// BasisTransformations.cpp:390-409).
__device__ inline void hierLine(double* v, int t, int s, bool inverse) {
    const int n = (1 << t) - 1;
    for (int kk = 0; kk < t - 1; ++kk) {
        const int k = inverse ? kk + 2 : t - kk;        // level being (de)hierarchized
        const int h = 1 << (t - k);
        for (int q = h; q <= n; q += 2 * h) {
            const double l = q - h >= 1 ? v[(q - h - 1) * s] : 0.0;
            const double r = q + h <= n ? v[(q + h - 1) * s] : 0.0;
            const double m = 0.5 * (l + r);
            v[(q - 1) * s] = inverse ? v[(q - 1) * s] + m : v[(q - 1) * s] - m;
        }
    }
}

// first factor of level t in the factor tables (levels 1.. t-1 before it), as facBase
__device__ inline int64_t facStart(int t) { return (int64_t(1) << t) - t - 1; }

// Solve M x = b on a line with the LU factors of its level (no pivoting).
__device__ inline void solveLine(double* v, int t, int s, const double* fac, int64_t facTot) {
    const int n = (1 << t) - 1;
    const double* l1 = fac + facStart(t);
    const double* l2 = l1 + facTot;
    const double* u0 = l1 + 2 * facTot;
    const double* u1 = l1 + 3 * facTot;
    const double* u2 = l1 + 4 * facTot;
    double y1 = 0.0, y2 = 0.0;   // y_{k-1}, y_{k-2}
    for (int k = 0; k < n; ++k) {
        const double y = v[k * s] - l1[k] * y1 - l2[k] * y2;
        v[k * s] = y;
        y2 = y1;
        y1 = y;
    }
    double x1 = 0.0, x2 = 0.0;   // x_{k+1}, x_{k+2}
    for (int k = n - 1; k >= 0; --k) {
        const double x = (v[k * s] - u1[k] * x1 - u2[k] * x2) / u0[k];
        v[k * s] = x;
        x2 = x1;
        x1 = x;
    }
}

// Coefficient of the basis function at line position j (1-based) in the nodal value at
// position r of a level-t line: the columns of PrewaveletMatrixHomogen::create
// (sgoperation/matrix.cpp:74-139).
__device__ inline double lineCoef(int t, int n, int j, int r) {
    const int o = r - j;
    if (j & 1) {
        if (o == 0) return (t > 1 && (j == 1 || j == n)) ? 0.9 : 1.0;
        if (o == 1 || o == -1) return -0.6;
        return 0.1;   // |o| == 2
    }
    if (o == 0) return 1.0;
    if (o == 1 || o == -1) return 0.5;
    return 0.0;
}

// v <- M v on a line (in place, old values kept in a 2-value window).
__device__ inline void applyLine(double* v, int t, int s) {
    const int n = (1 << t) - 1;
    if (n == 1) return;   // level 1: the identity
    double old2 = 0.0, old1 = 0.0;   // old v at r-2, r-1
    for (int r = 1; r <= n; ++r) {
        const double c0 = v[(r - 1) * s];
        double acc = 0.0;
        if (r - 2 >= 1) acc += lineCoef(t, n, r - 2, r) * old2;
        if (r - 1 >= 1) acc += lineCoef(t, n, r - 1, r) * old1;
        acc += lineCoef(t, n, r, r) * c0;
        if (r + 1 <= n) acc += lineCoef(t, n, r + 1, r) * v[r * s];
        if (r + 2 <= n) acc += lineCoef(t, n, r + 2, r) * v[(r + 1) * s];
        v[(r - 1) * s] = acc;
        old2 = old1;
        old1 = c0;
    }
}

// Line operations.  Operators along different dimensions commute, so a dehierarchization
// followed by the line solve (and the line-matrix product followed by the hierarchization)
// is done dimension by dimension in one pass over each line.
enum LineOp { L_DEHIER_SOLVE, L_APPLY_HIER };

// One pass of a line operation along every dimension of the g boxes of the workgroup
// (one thread per line).
__device__ void boxPass(double* box, const TrItem& it, int g, LineOp op, const double* fac, int64_t facTot) {
    for (int i = 0; i < it.d; ++i) {
        const int n = it.n[i];
        if (n > 1) {
            const int S = it.S[i];
            const int lines = it.total / n;
            for (int l = threadIdx.x; l < g * lines; l += blockDim.x) {
                const int j = l / lines;
                const int ll = l - j * lines;
                double* v = box + j * it.total + (ll % S) + (ll / S) * S * n;
                switch (op) {
                    case L_DEHIER_SOLVE:
                        hierLine(v, it.T[i], S, true);
                        solveLine(v, it.T[i], S, fac, facTot);
                        break;
                    case L_APPLY_HIER:
                        applyLine(v, it.T[i], S);
                        hierLine(v, it.T[i], S, false);
                        break;
                }
            }
        }
        __syncthreads();
    }
}

// box entry p is a point of depth exactly T (odd position in every dimension)
__device__ inline bool ownDepth(const TrItem& it, int p) {
    for (int i = 0; i < it.d; ++i) {
        const int k = p % it.n[i];
        p /= it.n[i];
        if (k & 1) return false;   // 0-based even = 1-based odd
    }
    return true;
}

// One workgroup per (depth, group of G instances): the g <= G boxes of the group sit
// one after another in LDS.
template <int MODE>
__global__ void __launch_bounds__(256) tr_box_kernel(const TrItem* __restrict__ items, const int32_t* __restrict__ list,
                                                      const int32_t* __restrict__ gidx, const double* __restrict__ fac,
                                                      int64_t facTot, const double* __restrict__ in,
                                                      double* __restrict__ out, double* __restrict__ scratch, int64_t n,
                                                      int64_t nml, int G, int batch) {
    extern __shared__ double box[];
    const TrItem it = items[list ? list[blockIdx.x] : blockIdx.x];
    const int64_t b0 = (int64_t)blockIdx.y * G;
    const int g = (int)min<int64_t>(G, batch - b0);
    const int tot = it.total;
    const int32_t* gi = gidx + it.box_off;
    for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
        const int j = p / tot;
        const int q = p - j * tot;
        if (MODE == TR_SYNTH && !ownDepth(it, q)) box[p] = 0.0;
        else box[p] = in[(b0 + j) * n + gi[q]];
    }
    __syncthreads();
    if (MODE == TR_LEVEL) {
        boxPass(box, it, g, L_DEHIER_SOLVE, fac, facTot);
        for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
            const int j = p / tot;
            const int q = p - j * tot;
            if (ownDepth(it, q)) out[(b0 + j) * n + gi[q]] = box[p];
            else box[p] = 0.0;
        }
        __syncthreads();
    }
    // TR_LEVEL / TR_SYNTH: hierarchical representation of the depth-T prewavelets on G_T
    boxPass(box, it, g, L_APPLY_HIER, fac, facTot);
    for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
        const int j = p / tot;
        const int q = p - j * tot;
        scratch[(b0 + j) * nml + it.box_off + q] = box[p];
    }
}

// one dimension of a hierarchization, point by point: out = in - (in[left pole] +
// in[right pole]) / 2 (0 on the boundary), the arithmetic of hierLine
__global__ void __launch_bounds__(256) tr_pole_kernel(const int32_t* __restrict__ poles, const double* __restrict__ in,
                                                      double* __restrict__ out, int64_t n) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int64_t b = blockIdx.y;
    const double* x = in + b * n;
    const int32_t lp = poles[2 * p], rp = poles[2 * p + 1];
    const double l = lp >= 0 ? x[lp] : 0.0;
    const double r = rp >= 0 ? x[rp] : 0.0;
    out[b * n + p] = x[p] - 0.5 * (l + r);
}

// one level of one dimension of a dehierarchization, in place: v[p] += (v[left pole] +
// v[right pole]) / 2 for the listed points (their poles are coarser along the dimension and
// final already), the arithmetic of hierLine(inverse)
__global__ void __launch_bounds__(256) tr_dehier_kernel(const int32_t* __restrict__ list, int64_t cnt,
                                                        const int32_t* __restrict__ poles, double* __restrict__ v,
                                                        int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cnt) return;
    double* x = v + (int64_t)blockIdx.y * n;
    const int32_t p = list[k];
    const int32_t lp = poles[2 * (int64_t)p], rp = poles[2 * (int64_t)p + 1];
    const double l = lp >= 0 ? x[lp] : 0.0;
    const double r = rp >= 0 ? x[rp] : 0.0;
    x[p] = x[p] + 0.5 * (l + r);
}

// v[rowDev[r]] = (sub ? v[rowDev[r]] : 0) -/+ sum of the row's scratch entries, in row order
__global__ void tr_gather_kernel(const int32_t* __restrict__ rowDev, const int64_t* __restrict__ rowPtr,
                                 const int32_t* __restrict__ col, int64_t nrows, const double* __restrict__ scratch,
                                 double* __restrict__ v, int64_t n, int64_t nml, int sub) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int64_t b = blockIdx.y;
    const double* z = scratch + b * nml;
    double s = 0.0;
    for (int64_t k = rowPtr[r]; k < rowPtr[r + 1]; ++k) s += z[col[k]];
    double* y = v + b * n + rowDev[r];
    *y = sub ? *y - s : s;
}

// ---- adaptive grids
// 1D (de)hierarchization of a line of a level-t box with entry status st: only active
// entries change (hanging points are not worked on, sgrid/sparseGrid.h:189-193), and
// entries that are not stored hold 0.
__device__ inline void hierLineMasked(double* v, const uint8_t* st, int t, int s, bool inverse) {
    const int n = (1 << t) - 1;
    for (int kk = 0; kk < t - 1; ++kk) {
        const int k = inverse ? kk + 2 : t - kk;
        const int h = 1 << (t - k);
        for (int q = h; q <= n; q += 2 * h) {
            if (st[(q - 1) * s] != TR_ACTIVE) continue;
            const double l = q - h >= 1 ? v[(q - h - 1) * s] : 0.0;
            const double r = q + h <= n ? v[(q + h - 1) * s] : 0.0;
            const double m = 0.5 * (l + r);
            v[(q - 1) * s] = inverse ? v[(q - 1) * s] + m : v[(q - 1) * s] - m;
        }
    }
}

// Solve M x = b in place on a line of n entries at stride s, M given by its unpivoted LU
// factors l1, l2, u0, u1, u2 (arrays facTot apart, f at this line's first l1 entry).
__device__ inline void solveBand(double* v, int n, int s, const double* f, int64_t facTot) {
    const double* l1 = f;
    const double* l2 = f + facTot;
    const double* u0 = f + 2 * facTot;
    const double* u1 = f + 3 * facTot;
    const double* u2 = f + 4 * facTot;
    double y1 = 0.0, y2 = 0.0;
    for (int k = 0; k < n; ++k) {
        const double y = v[k * s] - l1[k] * y1 - l2[k] * y2;
        v[k * s] = y;
        y2 = y1;
        y1 = y;
    }
    double x1 = 0.0, x2 = 0.0;
    for (int k = n - 1; k >= 0; --k) {
        const double x = (v[k * s] - u1[k] * x1 - u2[k] * x2) / u0[k];
        v[k * s] = x;
        x2 = x1;
        x1 = x;
    }
}

enum { TA_SEQ = 0, TA_SYNTH = 1 };

// Passes over whole lines of the g boxes (one thread per line): line op along dim i.
enum MaskedOp { M_DEHIER, M_HIER, M_APPLY };
__device__ void maskedPass(double* box0, int64_t boxStride, const uint8_t* st, const TrItem& it, int g, int i,
                           MaskedOp op) {
    const int n = it.n[i];
    if (n > 1) {
        const int S = it.S[i];
        const int lines = it.total / n;
        for (int l = threadIdx.x; l < g * lines; l += blockDim.x) {
            const int j = l / lines;
            const int ll = l - j * lines;
            const int64_t base = (ll % S) + (int64_t)(ll / S) * S * n;
            double* v = box0 + j * boxStride + base;
            switch (op) {
                case M_DEHIER: hierLineMasked(v, st + base, it.T[i], S, true); break;
                case M_HIER: hierLineMasked(v, st + base, it.T[i], S, false); break;
                case M_APPLY: applyLine(v, it.T[i], S); break;
            }
        }
    }
    __syncthreads();
}

// One workgroup per (depth, group of g instances).  TA_SEQ: one depth (`depth`) of
// calcPrewByNodal; src holds the nodal input (first depth) or aliases hier, the running
// hierarchical representation, which the kernel updates; out receives the prewavelets of
// the depth's active points.  TA_SYNTH: every depth (blockIdx.x), src holds prewavelets,
// the hierarchical representation of the depth's prewavelets goes to scratch.  The boxes
// sit in LDS (useLds) or in the scratch (boxTot entries per instance).
template <int MODE>
__global__ void __launch_bounds__(256) tr_adapt_kernel(
    const TrItem* __restrict__ items, int depth, const TrComp* __restrict__ comps, const int32_t* __restrict__ compOff,
    const int32_t* __restrict__ gidx, const uint8_t* __restrict__ bstat, const double* __restrict__ afac,
    int64_t afacTot, const double* src, double* hier, double* __restrict__ out, double* scratch, int64_t n,
    int64_t boxTot, int G, int batch, int first, int useLds) {
    extern __shared__ double lds[];
    const int t = MODE == TA_SEQ ? depth : (int)blockIdx.x;
    const TrItem it = items[t];
    const int64_t b0 = (int64_t)blockIdx.y * G;
    const int g = (int)min<int64_t>(G, batch - b0);
    const int tot = it.total;
    const int32_t* gi = gidx + it.box_off;
    const uint8_t* st = bstat + it.box_off;
    double* box0 = useLds ? lds : scratch + b0 * boxTot + it.box_off;
    const int64_t bs = useLds ? tot : boxTot;
    for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
        const int j = p / tot;
        const int q = p - j * tot;
        const int32_t dv = gi[q];
        double v = dv >= 0 ? src[(b0 + j) * n + dv] : 0.0;
        // CalcHierbyPrew scatters the prewavelets of every stored point of the depth,
        // hanging ones included (BasisTransformations.cpp:321-353)
        if (MODE == TA_SYNTH && !ownDepth(it, q)) v = 0.0;
        box0[j * bs + q] = v;
    }
    __syncthreads();
    if (MODE == TA_SEQ) {
        if (!first)
            for (int i = 0; i < it.d; ++i) maskedPass(box0, bs, st, it, g, i, M_DEHIER);
        // connected components, one after the other (their boxes may overlap)
        for (int c = compOff[t]; c < compOff[t + 1]; ++c) {
            const TrComp cc = comps[c];
            for (int i = 0; i < it.d; ++i) {
                int lines = 1;
                for (int k = 0; k < it.d; ++k)
                    if (k != i) lines *= cc.mx[k] - cc.mn[k] + 1;
                const int S = it.S[i];
                const int size = cc.size[i];
                for (int l = threadIdx.x; l < g * lines; l += blockDim.x) {
                    const int j = l / lines;
                    int ll = l - j * lines;
                    int64_t base = (int64_t)(cc.mn[i] - 1) * S;
                    for (int k = 0; k < it.d; ++k) {
                        if (k == i) continue;
                        const int ext = cc.mx[k] - cc.mn[k] + 1;
                        base += (int64_t)(cc.mn[k] + ll % ext - 1) * it.S[k];
                        ll /= ext;
                    }
                    double* v = box0 + j * bs + base;
                    solveBand(v, size, S, afac + cc.fac[i], afacTot);
                    // only active entries take the solution
                    for (int q = 0; q < size; ++q) {
                        const uint8_t e = st[base + (int64_t)q * S];
                        if (e == TR_ACTIVE) continue;
                        v[q * S] = e == TR_MISSING ? 0.0 : src[(b0 + j) * n + gi[base + (int64_t)q * S]];
                    }
                }
                __syncthreads();
            }
        }
        // prewavelets of the depth's active points (those of its hanging points are 0: the
        // output starts at 0 and only active points are written); everything else leaves
        // the box
        for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
            const int j = p / tot;
            const int q = p - j * tot;
            double& v = box0[j * bs + q];
            if (st[q] == TR_ACTIVE && ownDepth(it, q)) out[(b0 + j) * n + gi[q]] = v;
            else v = 0.0;
        }
        __syncthreads();
    }
    // CalcHierbyPrew: the prewavelets as nodal values on stored active entries, then
    // hierarchized over the active entries
    for (int i = 0; i < it.d; ++i) maskedPass(box0, bs, st, it, g, i, M_APPLY);
    for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
        const int q = p % tot;
        if (st[q] != TR_ACTIVE) box0[(p / tot) * bs + q] = 0.0;
    }
    __syncthreads();
    for (int i = 0; i < it.d; ++i) maskedPass(box0, bs, st, it, g, i, M_HIER);
    for (int p = threadIdx.x; p < g * tot; p += blockDim.x) {
        const int j = p / tot;
        const int q = p - j * tot;
        if (MODE == TA_SEQ) {
            if (st[q] == TR_ACTIVE) hier[(b0 + j) * n + gi[q]] -= box0[j * bs + q];
        } else if (useLds) {
            scratch[(b0 + j) * boxTot + it.box_off + q] = box0[j * bs + q];
        }
    }
}

// out[b] = sqrt(sum_i x_i y_i) for one instance per workgroup (fixed summation order)
__global__ void __launch_bounds__(256) tr_energy_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                                        int64_t n, double* __restrict__ out) {
    __shared__ double red[256];
    const int64_t b = blockIdx.x;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += x[b * n + i] * y[b * n + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[b] = sqrt(red[0]);
}

// out[2b] = max_i |x_i| over all stored points (L_infty, applications/norms.h:12-52),
// out[2b+1] = sqrt(sum x_i^2 / dofs) over the active points (L_2, norms.h:54-100); one
// workgroup per instance, fixed-order reductions
template <typename V>
__global__ void __launch_bounds__(256) vec_norms_kernel(const V* __restrict__ x, int64_t n,
                                                        const uint8_t* __restrict__ active, double dofs,
                                                        double* __restrict__ out) {
    __shared__ double mx[256], sq[256];
    const int64_t b = blockIdx.x;
    double m = 0.0, q = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = (double)x[b * n + i];
        m = fmax(m, fabs(v));
        if (!active || active[i]) q += v * v;
    }
    mx[threadIdx.x] = m;
    sq[threadIdx.x] = q;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            mx[threadIdx.x] = fmax(mx[threadIdx.x], mx[threadIdx.x + w]);
            sq[threadIdx.x] += sq[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * b] = mx[0];
        out[2 * b + 1] = sqrt(sq[0] / dofs);
    }
}

int boxThreads(int64_t vals) { return vals <= 128 ? 64 : (vals <= 512 ? 128 : 256); }

}  // namespace

void launch_tr_box(int mode, const TrItem* items, const int32_t* list, int nitems, const int32_t* gidx,
                   const double* fac, int64_t facTot, int64_t maxBox, const double* in, double* out, double* scratch,
                   int64_t n, int64_t nml, int batch, hipStream_t s) {
    if (nitems <= 0 || batch <= 0) return;
    // instances per workgroup: up to 32 KB of LDS (one box when a box needs more)
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>({16, batch, 4096 / std::max<int64_t>(maxBox, 1)}));
    const dim3 grid(nitems, (unsigned)((batch + G - 1) / G));
    const int thr = boxThreads((int64_t)G * maxBox);
    const size_t lds = (size_t)G * maxBox * sizeof(double);
    switch (mode) {
        case TR_LEVEL:
            tr_box_kernel<TR_LEVEL><<<grid, thr, lds, s>>>(items, list, gidx, fac, facTot, in, out, scratch, n, nml, G,
                                                            batch);
            break;
        default:
            tr_box_kernel<TR_SYNTH><<<grid, thr, lds, s>>>(items, list, gidx, fac, facTot, in, out, scratch, n, nml, G,
                                                            batch);
            break;
    }
}

void launch_tr_adapt(int mode, const TrItem* items, int depth, int nitems, int64_t maxBox, const TrComp* comps,
                     const int32_t* compOff, const int32_t* gidx, const uint8_t* bstat, const double* afac,
                     int64_t afacTot, const double* src, double* hier, double* out, double* scratch, int64_t n,
                     int64_t boxTot, int first, int64_t ldsLimit, int batch, hipStream_t s) {
    if (nitems <= 0 || batch <= 0) return;
    // instances per workgroup: up to 64 KB of LDS; one box per workgroup when it needs
    // more, in the scratch when it needs more than ldsLimit
    const int64_t boxBytes = std::max<int64_t>(maxBox, 1) * (int64_t)sizeof(double);
    int G = (int)std::max<int64_t>(1, std::min<int64_t>({16, batch, 65536 / boxBytes}));
    int useLds = 1;
    if ((int64_t)G * boxBytes > ldsLimit) {
        G = 1;
        useLds = boxBytes <= ldsLimit ? 1 : 0;
    }
    const dim3 grid(nitems, (unsigned)((batch + G - 1) / G));
    const size_t lds = useLds ? (size_t)G * boxBytes : 0;
    if (mode == TA_SEQ)
        tr_adapt_kernel<TA_SEQ><<<grid, 256, lds, s>>>(items, depth, comps, compOff, gidx, bstat, afac, afacTot, src,
                                                       hier, out, scratch, n, boxTot, G, batch, first, useLds);
    else
        tr_adapt_kernel<TA_SYNTH><<<grid, 256, lds, s>>>(items, depth, comps, compOff, gidx, bstat, afac, afacTot, src,
                                                         hier, out, scratch, n, boxTot, G, batch, first, useLds);
}

void launch_tr_pole(const int32_t* poles, const double* in, double* out, int64_t n, int batch, hipStream_t s) {
    if (n <= 0 || batch <= 0) return;
    tr_pole_kernel<<<dim3((unsigned)((n + 255) / 256), batch), 256, 0, s>>>(poles, in, out, n);
}

void launch_tr_dehier(const int32_t* list, int64_t cnt, const int32_t* poles, double* v, int64_t n, int batch,
                      hipStream_t s) {
    if (cnt <= 0 || batch <= 0) return;
    tr_dehier_kernel<<<dim3((unsigned)((cnt + 255) / 256), batch), 256, 0, s>>>(list, cnt, poles, v, n);
}

void launch_tr_gather(const int32_t* rowDev, const int64_t* rowPtr, const int32_t* col, int64_t nrows,
                      const double* scratch, double* v, int64_t n, int64_t nml, int sub, int batch, hipStream_t s) {
    if (nrows <= 0 || batch <= 0) return;
    const dim3 grid((unsigned)((nrows + 255) / 256), batch);
    tr_gather_kernel<<<grid, 256, 0, s>>>(rowDev, rowPtr, col, nrows, scratch, v, n, nml, sub);
}

template <typename V>
void launch_vec_norms(const V* x, int64_t n, const uint8_t* active, double dofs, double* out, int batch,
                      hipStream_t s) {
    if (batch <= 0) return;
    vec_norms_kernel<V><<<batch, 256, 0, s>>>(x, n, active, dofs, out);
}
template void launch_vec_norms<double>(const double*, int64_t, const uint8_t*, double, double*, int, hipStream_t);
template void launch_vec_norms<float>(const float*, int64_t, const uint8_t*, double, double*, int, hipStream_t);

void launch_tr_energy(const double* x, const double* y, int64_t n, double* out, int batch, hipStream_t s) {
    if (batch <= 0) return;
    tr_energy_kernel<<<batch, 256, 0, s>>>(x, y, n, out);
}

}  // namespace lsg
