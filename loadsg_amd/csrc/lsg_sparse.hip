// lsg_sparse.hip -- gfx950 kernels of the sparse path: grids with hanging nodes or
// partly filled depth blocks (locally adaptive sparse grids).
//
// The dense kernels work on full boxes G_T; here G_T is the list of stored points of
// depth <= T (MultiLevelAdaptiveSparseGrid, multilevelSparseGrid.h:13-143) and every
// stage is a gather over host-built tables (lsg_plan.cpp, buildSparse):
//   calcNodal            CSR rows per multilevel point        (MatrixVectorHomogen.cpp:43-74)
//   prolongation         (a, b) pairs per target point          (MatrixVectorHomogen.h:444-474)
//   stencil              cells of depth T+1 in the grid around each point, element
//                        matrices from the cell-centre coefficient (Stencil.h:187-246)
//   restriction          (self, left, right) per coarse point   (MatrixVectorHomogen.h:485-530)
//   ConvertToPrewavelet2 CSR rows per point of the target depth (MatrixVectorHomogen.h:540-562)
// The case/depth schedule, the gM binding and the wavefront levels are the dense
// path's (buildPlan), with one workgroup per (apply item, instance).  Passes of a
// workgroup communicate through global scratch, separated by a workgroup barrier.  Every sum has a fixed order.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "lsg_kernels.hpp"

namespace lsg {

namespace {

inline unsigned gridOf(int64_t n, int per = 256 * 4) {
    int64_t b = (n + per - 1) / per;
    if (b < 1) b = 1;
    if (b > 65535) b = 65535;
    return (unsigned)b;
}

// Passes of one workgroup hand data over through global memory.  All waves of a
// workgroup run on one CU, so the workgroup-scope release/acquire of the barrier
// makes the stores visible; a device-scope fence would write back and invalidate the
// whole L2 of the XCD at every pass.
__device__ inline void pass_barrier() {
    __syncthreads();
}

// coefficient a (stiffness terms) and c (mass terms) at the centre of the level-T cell
// with lower position lo (cell index lo[j] along dim j)
template <typename V>
__device__ inline void cell_coef(const V* __restrict__ coef, int64_t tabLen, int nA, int nC, int d,
                                 const int32_t* __restrict__ T, const int32_t* __restrict__ lo, V& aC, V& cC) {
    aC = V(0);
    cC = V(0);
    for (int m = 0; m < nA + nC; ++m) {
        V prod = V(1);
        for (int j = 0; j < d; ++j) prod *= coef[(int64_t)(m * d + j) * tabLen + ((1 << T[j]) - 2) + lo[j]];
        if (m < nA) aC += prod;
        else cC += prod;
    }
}

// element matrix entry between corners p and q (bit j = upper node in dim j) of a
// level-T cell: a * sum_i K1_i prod_{j!=i} M1_j + c * prod_j M1_j with
// K1 = 2^T [1 -1; -1 1], M1 = 2^-T [1/3 1/6; 1/6 1/3] (Poisson / HelmHoltz integration,
// Stencil.h:873-919, 1707-1752)
template <typename V>
__device__ inline V element(int d, const int32_t* __restrict__ T, unsigned p, unsigned q, V aC, V cC) {
    V mass = V(1);
    V stiff = V(0);
    for (int i = 0; i < d; ++i) {
        const bool same = ((p >> i) & 1u) == ((q >> i) & 1u);
        const V h = ldexp(V(1), -T[i]);
        const V m1 = (same ? V(1) / V(3) : V(1) / V(6)) * h;
        const V k1 = ldexp(same ? V(1) : V(-1), T[i]);       // = +-1 / h exactly
        // stiff = sum_i k1_i prod_{j != i} m1_j, accumulated with the mass product so far
        stiff = stiff * m1 + k1 * mass;
        mass *= m1;
    }
    return aC * stiff + cC * mass;
}

}  // namespace

// ---------------------------------------------------------- cell coefficients
// a, then c, at the centre of every cell of depth T+1 of every depth, per instance (a for
// all cells first, so an operator without a mass term reads one contiguous half); the
// coefficient does not change between matvecs, so the apply stage only loads these.
template <typename V>
__global__ void __launch_bounds__(256) sp_cellcoef_kernel(SparseDev s, int64_t nCells, int nA, int nC,
                                                          const V* __restrict__ coefTab, int64_t coefStride,
                                                          int64_t tabLen, V* __restrict__ cc, int64_t ccStride) {
    const int inst = blockIdx.y;
    const V* coef = coefTab + inst * coefStride;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nCells; c += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = s.nd;                 // depth of the cell: cellOff[t] <= c < cellOff[t + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s.cellOff[mid] <= c) lo = mid;
            else hi = mid;
        }
        V aC, cC;
        cell_coef<V>(coef, tabLen, nA, nC, s.d, s.depthT + (int64_t)lo * s.d, s.cellLo + c * s.d, aC, cC);
        cc[inst * ccStride + c] = aC;
        cc[inst * ccStride + nCells + c] = cC;
    }
}

// Instance slots of a workgroup that takes NI instances: slot j = first + k maps to
// map[j] (or j); slots past the batch, and instances that have stopped iterating, are off
// (they repeat an instance of the group for their loads and store nothing).  Returns
// whether any slot is on.
template <int NI>
__device__ inline bool group_slots(int first, int batch, const int* __restrict__ map,
                                   const CGState* __restrict__ st, int (&inst)[NI], bool (&on)[NI]) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        const int j = first + k;
        on[k] = j < batch;
        inst[k] = on[k] ? (map ? map[j] : j) : (map ? map[first] : first);
        if (on[k] && st && !st[inst[k]].active) on[k] = false;
        any = any || on[k];
    }
    return any;
}

// ------------------------------------------------------------------ calcNodal
// NI instances per thread: every (column, coefficient) entry is read once for all of them.
template <typename V, int NI>
__global__ void __launch_bounds__(256) sp_nodal_kernel(SparseDev s, const V* __restrict__ x, int64_t xStride,
                                                       V* __restrict__ nodal, int64_t nStride, int64_t nml, int batch,
                                                       const CGState* __restrict__ st, const int* __restrict__ map) {
    int inst[NI];
    bool on[NI];
    if (!group_slots<NI>((int)blockIdx.y * NI, batch, map, st, inst, on)) return;
    const V* xi[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) xi[k] = x + inst[k] * xStride;
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < nml; r += (int64_t)gridDim.x * blockDim.x) {
        V v[NI];
#pragma unroll
        for (int k = 0; k < NI; ++k) v[k] = V(0);
        const int64_t k1 = s.nodPtr[r + 1];
        for (int64_t e = s.nodPtr[r]; e < k1; ++e) {
            V w;
            if constexpr (std::is_same<V, float>::value) w = s.nodValF[e];
            else w = V(s.nodVal[e]);
            const int32_t c = s.nodCol[e];
#pragma unroll
            for (int k = 0; k < NI; ++k) v[k] = fma(w, xi[k][c], v[k]);
        }
#pragma unroll
        for (int k = 0; k < NI; ++k)
            if (on[k]) nodal[inst[k] * nStride + r] = v[k];
    }
}

// ---------------------------------------------------------------- prolongation
// One workgroup per (chain along dd, group of NI instances): box q = input(q) + I_dd(box q-1)
// (prolongation1D_inplace + addMultiLevelValues, MatrixVectorHomogen.h:444-474,997-1019).
template <typename V, int NI>
__global__ void __launch_bounds__(256) sp_prol_kernel(const ProlItem* __restrict__ items, SparseDev s,
                                                      const V* __restrict__ nodal, int64_t nStride,
                                                      V* __restrict__ cbuf, int64_t cStride, int batch,
                                                      const CGState* __restrict__ st, const int* __restrict__ map) {
    const int ng = (batch + NI - 1) / NI;
    int inst[NI];
    bool on[NI];
    if (!group_slots<NI>((int)(blockIdx.x % ng) * NI, batch, map, st, inst, on)) return;
    const ProlItem& it = items[blockIdx.x / ng];
    const int dd = it.dd;
    // box q reads the nodal buffer if bit q of from_nodal is set, else the case buffer;
    // box 0 keeps its values where they are (nothing coarser along dd)
    for (int q = 1; q < it.nt; ++q) {
        const int t = it.tidx[q];
        const int64_t size = s.boxSize[t];
        const int32_t* ab = s.prolAB + 2 * s.prolOff[(int64_t)t * s.d + dd];
        const bool inNodal = (it.from_nodal >> q) & 1;
        const V* in[NI];
        const V* prev[NI];
        V* out[NI];
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            const V* nodalI = nodal + inst[k] * nStride;
            const V* cbufI = cbuf + inst[k] * cStride;
            in[k] = (inNodal ? nodalI : cbufI) + it.in_off[q];
            out[k] = cbuf + inst[k] * cStride + it.out_off[q];
            prev[k] = q == 1 ? (((it.from_nodal & 1) ? nodalI : cbufI) + it.in_off[0]) : cbufI + it.out_off[q - 1];
        }
        for (int64_t i = threadIdx.x; i < size; i += blockDim.x) {
            const int32_t pa = ab[2 * i], pb = ab[2 * i + 1];
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                V add;
                if (pb == -2) add = prev[k][pa];
                else add = V(0.5) * ((pa >= 0 ? prev[k][pa] : V(0)) + (pb >= 0 ? prev[k][pb] : V(0)));
                const V v = in[k][i] + add;
                if (on[k]) out[k][i] = v;
            }
        }
        pass_barrier();
    }
}

// ---------------------------------------------------------------------- apply
// One workgroup per (case, depth) item and instance: stencil on G_T (+ gM), the
// restrictions with their gM exchanges, and the conversion into the part buffer.
// STORED: the element matrices are the caller's, one symmetric matrix per cell read from
// HBM (the stored local stiffness matrices of LocalStiffnessMatrices::applyLocalMatrix,
// localStiffnessMatrices/LocalStiffnessMatrices.cpp:506-531); D > 0 (d = D <=
// SP_CELL_MAXD): one thread loads the 2^D corner values of a cell once, forms the local
// product and stores it, then each point sums its entries of the cells around it; D = 0:
// each point gathers the corners of its cells.  Not STORED (D = 0 only; smaller d takes
// sp_apply_col_kernel): a(centre) times the Q1 stiffness plus c(centre) times the Q1 mass
// matrix, each point gathering the corners of its cells.
template <typename V, int D, bool STORED>
__global__ void __launch_bounds__(256) sp_apply_kernel(SparseApplyArgs<V> a) {
    const int inst = a.map ? a.map[blockIdx.x % a.batch] : (int)(blockIdx.x % a.batch);
    if (a.st && !a.st[inst].active) return;
    const ApplyItem& it = a.items[a.fam[blockIdx.x / a.batch].first];
    const SparseDev& s = a.sp;
    const int d = s.d;
    const int t = it.tidx;
    const int64_t size = s.boxSize[t];
    const V* u = (it.in_nodal ? a.nodal + inst * a.nStride : a.cbuf + inst * a.cStride) + it.in_off;
    V* A = a.zbuf + inst * a.zStride + it.z_off;
    V* B = A + size;
    V* wb = a.wbuf + inst * a.wStride;
    const V* cc = a.cellCoef + inst * a.ccStride;
    const unsigned nc = 1u << d;
    const bool mass = a.nC > 0;
    if constexpr (STORED) {
        const int64_t c0 = s.cellOff[t];
        const int64_t ncell = s.cellOff[t + 1] - c0;
        const V* E = a.cellMat + inst * a.matStride + c0;   // entry e of cell c: E[e * nCells + c]
        const int64_t es = a.nCells;
        const int64_t r0 = s.boxOff[t];
        const V* g = it.rd_off[0] >= 0 ? wb + it.rd_off[0] : nullptr;
        if constexpr (D > 0) {
            // one thread per cell: y = E u over the cell's corners (each entry read once,
            // consecutive cells of a wave read consecutive addresses), then each point sums
            // its entries of the cells around it
            constexpr int NC = 1 << D;
            V* Y = A + 2 * size;
            for (int64_t c = threadIdx.x; c < ncell; c += blockDim.x) {
                const int32_t* corner = s.cellCorner + (c0 + c) * NC;
                V uq[NC], y[NC];
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    const int32_t j = corner[q];
                    uq[q] = j >= 0 ? u[j] : V(0);
                    y[q] = V(0);
                }
                const V* Ec = E + c;
#pragma unroll
                for (int p = 0; p < NC; ++p)
#pragma unroll
                    for (int q = p; q < NC; ++q) {
                        const V m = Ec[(int64_t)cellEntry((unsigned)p, (unsigned)q, NC) * es];
                        y[p] = fma(m, uq[q], y[p]);
                        if (q != p) y[q] = fma(m, uq[p], y[q]);
                    }
#pragma unroll
                for (int p = 0; p < NC; ++p) Y[c * NC + p] = y[p];
            }
            pass_barrier();
            for (int64_t i = threadIdx.x; i < size; i += blockDim.x) {
                V z = V(0);
                const int64_t k1 = s.rcPtr[r0 + i + 1];
                for (int64_t k = s.rcPtr[r0 + i]; k < k1; ++k) z += Y[s.rcY[k]];
                A[i] = g ? z + g[i] : z;
            }
            pass_barrier();
        } else {
            // any d: each point forms its rows of the matrices of the cells around it
            for (int64_t i = threadIdx.x; i < size; i += blockDim.x) {
                V z = V(0);
                const int64_t k1 = s.rcPtr[r0 + i + 1];
                for (int64_t k = s.rcPtr[r0 + i]; k < k1; ++k) {
                    const int64_t cl = s.rcCell[k];
                    const unsigned cj = (unsigned)s.rcCorner[k];
                    const int32_t* corner = s.cellCorner + (c0 + cl) * nc;
                    for (unsigned q = 0; q < nc; ++q) {
                        const int32_t j = corner[q];
                        if (j < 0) continue;
                        const int e = q >= cj ? cellEntry(cj, q, nc) : cellEntry(q, cj, nc);
                        z = fma(E[(int64_t)e * es + cl], u[j], z);
                    }
                }
                A[i] = g ? z + g[i] : z;
            }
            pass_barrier();
        }
    } else {
    // element matrix of a level-T cell, split into its stiffness and mass parts: the
    // entry between corners p and q depends only on the mask p ^ q
    __shared__ V eS[1 << MAXD], eM[1 << MAXD];
    for (unsigned m = threadIdx.x; m < nc; m += blockDim.x) {
        eS[m] = element<V>(d, it.T, 0u, m, V(1), V(0));
        eM[m] = element<V>(d, it.T, 0u, m, V(0), V(1));
    }
    __syncthreads();
    // stencil: z_i = sum over the cells around point i of a_cell (K u)_i + c_cell (M u)_i
    {
        const int64_t r0 = s.boxOff[t];
        const int64_t c0 = s.cellOff[t];
        const V* g = it.rd_off[0] >= 0 ? wb + it.rd_off[0] : nullptr;
        for (int64_t i = threadIdx.x; i < size; i += blockDim.x) {
            V z = V(0);
            const int64_t k1 = s.rcPtr[r0 + i + 1];
            for (int64_t k = s.rcPtr[r0 + i]; k < k1; ++k) {
                const int64_t cell = c0 + s.rcCell[k];
                const unsigned cj = (unsigned)s.rcCorner[k];
                const int32_t* corner = s.cellCorner + cell * nc;
                V zs = V(0), zm = V(0);
                for (unsigned q = 0; q < nc; ++q) {
                    const int32_t j = corner[q];
                    if (j < 0) continue;
                    const V uj = u[j];
                    zs = fma(eS[cj ^ q], uj, zs);
                    if (mass) zm = fma(eM[cj ^ q], uj, zm);
                }
                z = fma(cc[cell], zs, z);
                if (mass) z = fma(cc[a.nCells + cell], zm, z);
            }
            A[i] = g ? z + g[i] : z;
        }
        pass_barrier();
    }
    }
    // restrictions along the restriction dims (ascending) with the gM exchange
    int64_t cur = size;
    for (int j = 1; j <= it.k; ++j) {
        const int dd = it.r[j - 1];
        const int tf = it.ctidx[j - 1], tc = it.ctidx[j];
        const int64_t nco = s.boxSize[tc];
        const int32_t* slr = s.restSLR + 3 * s.restOff[(int64_t)tf * d + dd];
        const bool w = it.w_off[j] >= 0, rd = it.rd_off[j] >= 0;
        V* wd = wb + it.w_off[j];
        const V* g = wb + it.rd_off[j];
        for (int64_t i = threadIdx.x; i < nco; i += blockDim.x) {
            const int32_t sf = slr[3 * i], l = slr[3 * i + 1], r = slr[3 * i + 2];
            const V v = A[sf] + V(0.5) * ((l >= 0 ? A[l] : V(0)) + (r >= 0 ? A[r] : V(0)));
            if (w) wd[i] = v;
            B[i] = rd ? v + g[i] : v;
        }
        pass_barrier();
        V* tmp = A; A = B; B = tmp;
        cur = nco;
    }
    (void)cur;
    // prewavelet conversion of the points of the target depth
    {
        const int tc = it.ctidx[it.k];
        const int64_t b0 = s.blkOff[tc], nb = s.blkSize[tc];
        V* out = a.part + inst * a.pStride + it.out_off;
        for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) {
            V v = V(0);
            for (int64_t k = s.convPtr[b0 + i]; k < s.convPtr[b0 + i + 1]; ++k) v += V(s.convVal[k]) * A[s.convCol[k]];
            out[i] = v;
        }
    }
}

// The cell-wise apply (D = d <= SP_CELL_MAXD, element matrices from the cell coefficients)
// with the stencil accumulated point by point: the cells of depth T+1 are taken colour by
// colour (parity of their position; cells of one colour share no corner), each thread forms
// a cell's local product and adds it to the cell's corners, so every point receives its
// cells' entries in a fixed order (gM value first, then colour 0, 1, ...) without an
// intermediate per-cell array.  LDSM: the point values (the box and, after the first
// restriction, the coarser box) live in LDS, ldsStride values per instance; otherwise in
// zbuf.  NI instances per workgroup share every index-table load (cell corners, restriction
// triples, conversion columns and weights); slots past the batch (the last group) repeat the
// group's first instance and store nothing.
template <typename V, int D, int NI, bool LDSM, int NT, typename IX>
__global__ void __launch_bounds__(NT) sp_apply_col_kernel(SparseApplyArgs<V> a) {
    extern __shared__ char smem_raw[];
    const int ng = (a.batch + NI - 1) / NI;
    const int grp = (int)(blockIdx.x % ng);
    const ApplyItem& it = a.items[a.fam[blockIdx.x / ng].first];
    const SparseDev& s = a.sp;
    // index tables: 16-bit copies when every box and cell list of the grid fits them
    const IX* cellCornerT;
    const IX* cellColT;
    const IX* restSLRT;
    const IX* convColT;
    if constexpr (sizeof(IX) == 2) {
        cellCornerT = s.cellCorner16; cellColT = s.cellCol16; restSLRT = s.restSLR16; convColT = s.convCol16;
    } else {
        cellCornerT = s.cellCorner; cellColT = s.cellCol; restSLRT = s.restSLR; convColT = s.convCol;
    }
    int inst[NI];
    bool on[NI];
    if (!group_slots<NI>(grp * NI, a.batch, a.map, a.st, inst, on)) return;
    const int t = it.tidx;
    const int64_t size = s.boxSize[t];
    constexpr int NC = 1 << D;
    const V* u[NI];
    V* A[NI];
    V* B[NI];
    const V* cc[NI];
    V* wb[NI];
#pragma unroll
    for (int k = 0; k < NI; ++k) {
        u[k] = (it.in_nodal ? a.nodal + inst[k] * a.nStride : a.cbuf + inst[k] * a.cStride) + it.in_off;
        if constexpr (LDSM) A[k] = reinterpret_cast<V*>(smem_raw) + k * a.ldsStride;
        else A[k] = a.zbuf + inst[k] * a.zStride + it.z_off;
        B[k] = A[k] + size;
        cc[k] = a.cellCoef + inst[k] * a.ccStride;
        wb[k] = a.wbuf + inst[k] * a.wStride;
    }
    const bool mass = a.nC > 0;
    // the gM value (or 0) starts every point's sum
    {
        const bool rd = it.rd_off[0] >= 0;
        for (int64_t i = threadIdx.x; i < size; i += NT)
#pragma unroll
            for (int k = 0; k < NI; ++k)
                if (on[k]) A[k][i] = rd ? wb[k][it.rd_off[0] + i] : V(0);
    }
    V eS[NC], eM[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        eS[m] = element<V>(D, it.T, 0u, (unsigned)m, V(1), V(0));
        eM[m] = element<V>(D, it.T, 0u, (unsigned)m, V(0), V(1));
    }
    __syncthreads();
    const int64_t c0 = s.cellOff[t];
    const int64_t* colOff = s.cellColOff + (int64_t)t * (NC + 1);
    for (int col = 0; col < NC; ++col) {
        const int64_t e1 = colOff[col + 1];
        for (int64_t e = colOff[col] + threadIdx.x; e < e1; e += NT) {
            const int64_t c = c0 + cellColT[e];
            const IX* corner = cellCornerT + c * NC;
            int32_t jq[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) jq[q] = corner[q];
            V uq[NI][NC];
            V aC[NI], cC[NI];
#pragma unroll
            for (int k = 0; k < NI; ++k) {
#pragma unroll
                for (int q = 0; q < NC; ++q) uq[k][q] = jq[q] >= 0 ? u[k][jq[q]] : V(0);
                aC[k] = cc[k][c];
                cC[k] = mass ? cc[k][a.nCells + c] : V(0);
            }
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                if (!on[k]) continue;
#pragma unroll
                for (int p = 0; p < NC; ++p) {
                    V acc = V(0);
#pragma unroll
                    for (int q = 0; q < NC; ++q) acc = fma(eS[p ^ q], uq[k][q], acc);
                    V y = aC[k] * acc;
                    if (mass) {
                        V accm = V(0);
#pragma unroll
                        for (int q = 0; q < NC; ++q) accm = fma(eM[p ^ q], uq[k][q], accm);
                        y = fma(cC[k], accm, y);
                    }
                    if (jq[p] >= 0) A[k][jq[p]] += y;
                }
            }
        }
        __syncthreads();
    }
    // restrictions along the restriction dims (ascending) with the gM exchange
    for (int j = 1; j <= it.k; ++j) {
        const int dd = it.r[j - 1];
        const int tf = it.ctidx[j - 1], tc = it.ctidx[j];
        const int64_t nco = s.boxSize[tc];
        const IX* slr = restSLRT + 3 * s.restOff[(int64_t)tf * s.d + dd];
        const bool w = it.w_off[j] >= 0, rd = it.rd_off[j] >= 0;
        for (int64_t i = threadIdx.x; i < nco; i += NT) {
            const int32_t sf = slr[3 * i], l = slr[3 * i + 1], r = slr[3 * i + 2];
#pragma unroll
            for (int k = 0; k < NI; ++k) {
                if (!on[k]) continue;
                const V v = A[k][sf] + V(0.5) * ((l >= 0 ? A[k][l] : V(0)) + (r >= 0 ? A[k][r] : V(0)));
                if (w) wb[k][it.w_off[j] + i] = v;
                B[k][i] = rd ? v + wb[k][it.rd_off[j] + i] : v;
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NI; ++k) {
            V* tmp = A[k]; A[k] = B[k]; B[k] = tmp;
        }
    }
    // prewavelet conversion of the points of the target depth
    {
        const int tc = it.ctidx[it.k];
        const int64_t b0 = s.blkOff[tc], nb = s.blkSize[tc];
        for (int64_t i = threadIdx.x; i < nb; i += NT) {
            V v[NI];
#pragma unroll
            for (int k = 0; k < NI; ++k) v[k] = V(0);
            const int64_t e1 = s.convPtr[b0 + i + 1];
            for (int64_t e = s.convPtr[b0 + i]; e < e1; ++e) {
                V wgt;
                if constexpr (std::is_same<V, float>::value) wgt = s.convValF[e];
                else wgt = V(s.convVal[e]);
                const int32_t col = convColT[e];
#pragma unroll
                for (int k = 0; k < NI; ++k) v[k] += wgt * A[k][col];
            }
#pragma unroll
            for (int k = 0; k < NI; ++k)
                if (on[k]) (a.part + inst[k] * a.pStride + it.out_off)[i] = v[k];
        }
    }
}

// ------------------------------------------------------------------- diagonal
// z_p = psi_p^T A_T psi_p (Preconditioning::precond3, applications/cg_method.h:322-451,
// with applyStencilLocal over the cells of depth T+1 in the grid): for every stored
// point J of p's prewavelet stencil, the cells around J, their corners q and the
// stencil coefficient psi_p(q) from the position offset.  Hanging nodes get 1.
// With stored element matrices (cellMat) the entries come from them, as in
// Preconditioning::precondLocalStiffnessMatrix (cg_method.h:544-700).
template <typename V>
__global__ void __launch_bounds__(256) sp_diag_kernel(SparseDev s, const uint8_t* __restrict__ activeDev, int n,
                                                      const V* __restrict__ cellCoef, int64_t ccStride,
                                                      const V* __restrict__ cellMat, int64_t matStride, int64_t nCells,
                                                      V* __restrict__ z, int64_t zStride) {
    const int inst = blockIdx.y;
    const V* cc = cellCoef + inst * ccStride;
    const V* E = cellMat ? cellMat + inst * matStride : nullptr;
    const int d = s.d;
    const unsigned nc = 1u << d;
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (!activeDev[p]) {
            z[inst * zStride + p] = V(1);
            continue;
        }
        const int t = s.tOfDev[p];
        const int32_t* T = s.depthT + (int64_t)t * d;
        const int32_t* pp = s.pos + p * d;
        const int64_t r0 = s.boxOff[t], c0 = s.cellOff[t];
        V total = V(0);
        for (int64_t k = s.convPtr[p]; k < s.convPtr[p + 1]; ++k) {
            const int32_t J = s.convCol[k];
            V row = V(0);
            for (int64_t e = s.rcPtr[r0 + J]; e < s.rcPtr[r0 + J + 1]; ++e) {
                const int64_t cell = c0 + s.rcCell[e];
                const unsigned cj = (unsigned)s.rcCorner[e];
                const int32_t* lo = s.cellLo + cell * d;
                const V aC = E ? V(0) : cc[cell], cC = E ? V(0) : cc[nCells + cell];
                const int32_t* corner = s.cellCorner + cell * nc;
                for (unsigned q = 0; q < nc; ++q) {
                    if (corner[q] < 0) continue;
                    // psi_p at corner q: product of the 1D stencil weights by offset
                    V psi = V(1);
                    for (int i = 0; i < d; ++i) {
                        const int off = lo[i] + (int)((q >> i) & 1u) - pp[i];
                        const int n1 = (1 << T[i]) - 1;
                        V c1;
                        if (off == 0) c1 = (T[i] >= 2 && (pp[i] == 1 || pp[i] == n1)) ? V(0.9) : V(1);
                        else if (off == 1 || off == -1) c1 = V(-0.6);
                        else if (off == 2 || off == -2) c1 = V(0.1);
                        else c1 = V(0);
                        psi *= c1;
                    }
                    if (psi == V(0)) continue;
                    if (E) row += E[(int64_t)(q >= cj ? cellEntry(cj, q, nc) : cellEntry(q, cj, nc)) * nCells + cell] * psi;
                    else row += element<V>(d, T, cj, q, aC, cC) * psi;
                }
            }
            total += V(s.convVal[k]) * row;
        }
        z[inst * zStride + p] = total;
    }
}

// -------------------------------------------------------------------- launchers
template <typename V>
void launch_sp_cellcoef(const SparseDev& s, int64_t nCells, int nA, int nC, const V* coefTab, int64_t coefStride,
                        int64_t tabLen, V* cc, int64_t ccStride, int batch, hipStream_t str) {
    if (nCells == 0) return;
    hipLaunchKernelGGL(sp_cellcoef_kernel<V>, dim3(gridOf(nCells, 256), batch), dim3(256), 0, str, s, nCells, nA, nC,
                       coefTab, coefStride, tabLen, cc, ccStride);
}
template <typename V>
void launch_sp_nodal(const SparseDev& s, const V* x, int64_t xStride, V* nodal, int64_t nStride, int64_t nml,
                     int batch, const CGState* st, hipStream_t str, const int* map) {
    // four instances per thread from batch 2 on
    if (batch > 1)
        hipLaunchKernelGGL((sp_nodal_kernel<V, 4>), dim3(gridOf(nml), (batch + 3) / 4), dim3(256), 0, str, s, x,
                           xStride, nodal, nStride, nml, batch, st, map);
    else
        hipLaunchKernelGGL((sp_nodal_kernel<V, 1>), dim3(gridOf(nml), batch), dim3(256), 0, str, s, x, xStride,
                           nodal, nStride, nml, batch, st, map);
}
template <typename V>
void launch_sp_prol(const ProlItem* items, int nitems, const SparseDev& s, const V* nodal, int64_t nStride, V* cbuf,
                    int64_t cStride, int batch, const CGState* st, hipStream_t str, const int* map) {
    if (nitems == 0) return;
    // one instance per workgroup: the chains are serial box by box, and grouping instances
    // cut the workgroups in flight more than it saved in table loads (level-11 adaptive grid,
    // batch 64: 1.40 ms with four instances per workgroup, 0.83 ms with one)
    hipLaunchKernelGGL((sp_prol_kernel<V, 1>), dim3((unsigned)nitems * (unsigned)batch), dim3(256), 0, str, items, s,
                       nodal, nStride, cbuf, cStride, batch, st, map);
}
template <typename V>
void launch_sp_apply(const SparseApplyArgs<V>& a, int nitems, hipStream_t str, int cls) {
    if (nitems == 0) return;
    const dim3 grid((unsigned)nitems * (unsigned)a.batch);
    static_assert(SP_CELL_MAXD == 4, "one instantiation per cell-wise dimension");
    if (a.cellMat) {
        switch (a.sp.d) {
            case 1: hipLaunchKernelGGL((sp_apply_kernel<V, 1, true>), grid, dim3(256), 0, str, a); break;
            case 2: hipLaunchKernelGGL((sp_apply_kernel<V, 2, true>), grid, dim3(256), 0, str, a); break;
            case 3: hipLaunchKernelGGL((sp_apply_kernel<V, 3, true>), grid, dim3(256), 0, str, a); break;
            case 4: hipLaunchKernelGGL((sp_apply_kernel<V, 4, true>), grid, dim3(256), 0, str, a); break;
            default: hipLaunchKernelGGL((sp_apply_kernel<V, 0, true>), grid, dim3(256), 0, str, a); break;
        }
        return;
    }
    if (a.sp.d > SP_CELL_MAXD) {
        hipLaunchKernelGGL((sp_apply_kernel<V, 0, false>), grid, dim3(256), 0, str, a);
        return;
    }
    // cell-wise kernels by scratch class: SP_SMALL (LDS, 4 instances per workgroup from
    // batch 2 on), SP_MEDIUM (LDS, one instance, 256 threads), SP_LARGE (LDS, one instance,
    // 1024 threads), SP_GLOBAL (zbuf, 4 instances per workgroup)
    const bool grp = a.batch > 1 && (cls == SP_SMALL || cls == SP_GLOBAL);
    const dim3 gg((unsigned)nitems * (unsigned)(grp ? (a.batch + 3) / 4 : a.batch));
    const size_t lds = (size_t)a.ldsStride * sizeof(V) * (grp ? 4 : 1);
#define LSG_SP_COL(DD, IX)                                                                                        \
    switch (cls) {                                                                                                \
        case SP_SMALL:                                                                                            \
            if (grp) hipLaunchKernelGGL((sp_apply_col_kernel<V, DD, 4, true, 256, IX>), gg, dim3(256), lds, str, a); \
            else hipLaunchKernelGGL((sp_apply_col_kernel<V, DD, 1, true, 256, IX>), gg, dim3(256), lds, str, a);     \
            break;                                                                                                \
        case SP_MEDIUM:                                                                                           \
            hipLaunchKernelGGL((sp_apply_col_kernel<V, DD, 1, true, 256, IX>), gg, dim3(256), lds, str, a);          \
            break;                                                                                                \
        case SP_LARGE:                                                                                            \
            hipLaunchKernelGGL((sp_apply_col_kernel<V, DD, 1, true, 1024, IX>), gg, dim3(1024), lds, str, a);        \
            break;                                                                                                \
        default:                                                                                                  \
            if (grp) hipLaunchKernelGGL((sp_apply_col_kernel<V, DD, 4, false, 256, IX>), gg, dim3(256), 0, str, a);  \
            else hipLaunchKernelGGL((sp_apply_col_kernel<V, DD, 1, false, 256, IX>), gg, dim3(256), 0, str, a);      \
            break;                                                                                                \
    }
    const bool ix16 = a.sp.cellCorner16 != nullptr;
    switch (a.sp.d) {
        case 1: if (ix16) { LSG_SP_COL(1, int16_t) } else { LSG_SP_COL(1, int32_t) } break;
        case 2: if (ix16) { LSG_SP_COL(2, int16_t) } else { LSG_SP_COL(2, int32_t) } break;
        case 3: if (ix16) { LSG_SP_COL(3, int16_t) } else { LSG_SP_COL(3, int32_t) } break;
        default: if (ix16) { LSG_SP_COL(4, int16_t) } else { LSG_SP_COL(4, int32_t) } break;
    }
#undef LSG_SP_COL
}
template <typename V>
void launch_sp_diag(const SparseDev& s, const uint8_t* activeDev, int n, const V* cellCoef, int64_t ccStride,
                    const V* cellMat, int64_t matStride, int64_t nCells, V* z, int64_t zStride, int batch,
                    hipStream_t str) {
    hipLaunchKernelGGL(sp_diag_kernel<V>, dim3(gridOf(n, 256), batch), dim3(256), 0, str, s, activeDev, n, cellCoef,
                       ccStride, cellMat, matStride, nCells, z, zStride);
}

#define LSG_SP_INST(V)                                                                                              \
    template void launch_sp_cellcoef<V>(const SparseDev&, int64_t, int, int, const V*, int64_t, int64_t, V*, int64_t, \
                                        int, hipStream_t);                                                          \
    template void launch_sp_nodal<V>(const SparseDev&, const V*, int64_t, V*, int64_t, int64_t, int, const CGState*, \
                                     hipStream_t, const int*);                                                      \
    template void launch_sp_prol<V>(const ProlItem*, int, const SparseDev&, const V*, int64_t, V*, int64_t, int,    \
                                    const CGState*, hipStream_t, const int*);                                       \
    template void launch_sp_apply<V>(const SparseApplyArgs<V>&, int, hipStream_t, int);                                  \
    template void launch_sp_diag<V>(const SparseDev&, const uint8_t*, int, const V*, int64_t, const V*, int64_t,     \
                                    int64_t, V*, int64_t, int, hipStream_t);
LSG_SP_INST(double)
LSG_SP_INST(float)
#undef LSG_SP_INST

}  // namespace lsg
