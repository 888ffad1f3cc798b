// lsg_kernels.hip -- gfx950 kernels of the matrix-free prewavelet matvec and CG.
//
// One workgroup owns one work item (a depth box, a prolongation family, or a
// (case, depth) apply item) of one batch instance and keeps the whole box in LDS;
// every stage is a sequence of 1D sweeps along one dimension of the box.  Boxes at
// the target shapes are small (d=5 level 7: at most 1323 points), so the global
// traffic of a stage is its algorithmic input and output read/written once,
// coalesced, and the sweeps run from LDS.  Reductions are deterministic
// (fixed-order wave shuffles + fixed-order partial sums), so results are
// bit-reproducible for a given launch shape.
#include <hip/hip_runtime.h>
#include <type_traits>

#include "lsg_kernels.hpp"

namespace lsg {

// ---------------------------------------------------------------- box shapes
// Box shape shared by the workgroup (in LDS; thread 0 updates it between barriers).
// Integer division by an extent or stride uses the float reciprocal: exact for
// dividends below 2^16 (boxes are limited to 65535 points by the host), since the
// quotient's fractional part (r + 0.5)/n is at least 0.5/n away from an integer.
struct Shape {
    int d;
    int total;
    int n[MAXD];      // extent per dim
    int S[MAXD];      // stride per dim (dim 0 fastest)
    float rn[MAXD];   // 1/n
    float rS[MAXD];   // 1/S
};

// The per-launch item tables are read-only for the kernels that use them: read them
// through the constant address space so that the uniform loads become scalar loads
// (their waits do not block the vector-memory loads in flight).
#define LSG_CONST __attribute__((address_space(4)))
template <typename T>
__device__ inline const LSG_CONST T* cptr(const T* p) { return (const LSG_CONST T*)p; }

// floor(p / n) for 0 <= p < 2^16 from r = rcp(n): (p + 0.5) r rounded once (fma with the
// exact 0.5 r), truncated by the conversion
__device__ inline int fdiv(int p, float r) { return (int)__builtin_fmaf(float(p), r, 0.5f * r); }
// products of box indices and extents (all below 2^16): the full-rate 24-bit multiply
__device__ inline int mul24(int a, int b) { return __umul24((unsigned)a, (unsigned)b); }
// Reciprocal of a positive integer for fdiv: the hardware reciprocal (1 ulp) keeps the
// quotient exact, since its relative error times (p + 0.5)/n stays below 2^-6/n,
// under the 0.5/n distance of (p + 0.5)/n from an integer.
__device__ inline float rcpf(int n) { return __builtin_amdgcn_rcpf(float(n)); }
// 1 / (6 * 2^t), exactly as the correctly rounded quotient (scaling by 2^-t is exact)
template <typename V>
__device__ inline V inv6pow2(int t) { return ldexp(V(1.0) / V(6.0), -t); }

__device__ inline void shape_restage(Shape& sh) {
    int s = 1;
    for (int q = 0; q < sh.d; ++q) {
        sh.S[q] = s;
        sh.rS[q] = rcpf(s);
        sh.rn[q] = rcpf(sh.n[q]);
        s *= sh.n[q];
    }
    sh.total = s;
}
template <typename TP>
__device__ inline void shape_box(Shape& sh, TP T, int d) {
    if (threadIdx.x == 0) {
        sh.d = d;
        for (int i = 0; i < d; ++i) sh.n[i] = (1 << T[i]) - 1;
        shape_restage(sh);
    }
    __syncthreads();
}
__device__ inline void shape_set_extent(Shape& sh, int i, int n) {
    __syncthreads();
    if (threadIdx.x == 0) {
        sh.n[i] = n;
        shape_restage(sh);
    }
    __syncthreads();
}
// coordinate of point p along dim i
__device__ inline int coord(const Shape& sh, int p, int i) {
    int q = fdiv(p, sh.rS[i]);
    return q - fdiv(q, sh.rn[i]) * sh.n[i];
}

template <typename V>
__device__ inline V cCenter(int t, int k, int n) {      // 0.9 next to the boundary (index.cc:312-316)
    return (t >= 2 && (k == 1 || k == n)) ? V(0.9) : V(1.0);
}

template <typename V>
__device__ inline void copy_in(V* __restrict__ dst, const V* __restrict__ src, int cnt) {
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) dst[p] = src[p];
}

// -------------------------------------------------------- point sweeps
// One output point per thread and step; all reads of a step are independent LDS
// loads.  (lo, k, hi) decomposition of an output index along the swept dim uses the
// exact reciprocal division above.
struct Split {
    int lo, k, hi;
};
__device__ inline Split split(int p, int Slo, float rSlo, int n, float rn) {
    Split s;
    int q = fdiv(p, rSlo);
    s.lo = p - mul24(q, Slo);
    s.hi = fdiv(q, rn);
    s.k = q - mul24(s.hi, n);
    return s;
}

// 1D prewavelet -> nodal along one dim (calcNodal, MatrixVectorHomogen.cpp:43-74, with
// the 5^d stencil of nextFiveP, indices/index.cc:252-327): m = 2^(t-1) coefficients at
// the odd positions -> n = 2^t - 1 nodal values.
template <typename V>
__device__ inline void expand_dim(const V* __restrict__ src, V* __restrict__ dst, int Slo, int m, int n, int outCnt,
                                  int t) {
    const float rSlo = rcpf(Slo), rn = rcpf(n);
    for (int p = threadIdx.x; p < outCnt; p += blockDim.x) {
        Split s = split(p, Slo, rSlo, n, rn);
        const V* b = src + s.lo + mul24(s.hi, Slo * m);
        const int k = s.k + 1;      // 1-based position
        V v;
        if (k & 1) {
            const int j = s.k >> 1;
            const V* bj = b + mul24(j, Slo);
            v = cCenter<V>(t, k, n) * bj[0];
            if (j >= 1) v += V(0.1) * bj[-Slo];
            if (j + 1 < m) v += V(0.1) * bj[Slo];
        } else {
            const int j = (k >> 1) - 1;
            const V* bj = b + mul24(j, Slo);
            v = V(-0.6) * (bj[0] + bj[Slo]);
        }
        dst[p] = v;
    }
}

// ConvertToPrewavelet2 along one dim (MatrixVectorHomogen.h:540-562): n nodal-dual
// values -> m = (n+1)/2 prewavelet values at the odd positions.
template <typename V>
__device__ inline void contract_dim(const V* __restrict__ src, V* __restrict__ dst, int Slo, int n, int outCnt,
                                    int t) {
    const int m = (n + 1) >> 1;
    const float rSlo = rcpf(Slo), rm = rcpf(m);
    for (int p = threadIdx.x; p < outCnt; p += blockDim.x) {
        Split s = split(p, Slo, rSlo, m, rm);
        const V* b = src + s.lo + mul24(s.hi, Slo * n);
        const int k = 2 * s.k + 1;            // 1-based position, 0-based index k-1
        const V* bk = b + mul24(k - 1, Slo);
        V v = cCenter<V>(t, k, n) * bk[0];
        if (k - 1 >= 1) v += V(-0.6) * bk[-Slo] + V(0.1) * bk[-2 * Slo];
        if (k + 1 <= n) v += V(-0.6) * bk[Slo] + V(0.1) * bk[2 * Slo];
        dst[p] = v;
    }
}

// restriction1D_inverted_inplace along r (MatrixVectorHomogen.h:485-530): coarse
// position j gets z(2j) + 0.5 (z(2j-1) + z(2j+1)); src box -> dst box.
template <typename V, typename SH>
__device__ inline void restrict_dim(const V* __restrict__ src, V* __restrict__ dst, const SH& sh, int r) {
    const int nf = sh.n[r], nc = (nf - 1) >> 1, S = sh.S[r];
    const int outCnt = sh.total / nf * nc;
    const float rS = sh.rS[r], rnc = rcpf(nc);
    for (int p = threadIdx.x; p < outCnt; p += blockDim.x) {
        Split s = split(p, S, rS, nc, rnc);
        const V* b = src + s.lo + mul24(s.hi, S * nf) + mul24(2 * s.k + 1, S);
        dst[p] = b[0] + V(0.5) * (b[-S] + b[S]);
    }
}

// prolongation1D_inplace + addMultiLevelValues along fd (MatrixVectorHomogen.h:444-474,
// 997-1019): cur (fine along fd) += I_fd(prev) (coarse along fd); sh = fine shape.
template <typename V, typename SH>
__device__ inline void prolong_dim(V* __restrict__ cur, const V* __restrict__ prev, const SH& sh, int fd) {
    const int nf = sh.n[fd], nc = (nf - 1) >> 1, S = sh.S[fd];
    const float rS = sh.rS[fd], rnf = sh.rn[fd];
    for (int p = threadIdx.x; p < sh.total; p += blockDim.x) {
        Split s = split(p, S, rS, nf, rnf);
        const V* c = prev + s.lo + mul24(s.hi, S * nc);
        const int k = s.k + 1;                // fine position
        V add;
        const V* cb = c + mul24((k - 1) >> 1, S);     // coarse position (k-1)/2 + 1 (even k: k/2)
        if (k & 1) {
            const int cl = (k - 1) >> 1, cr = (k + 1) >> 1;   // coarse positions 0..nc+1
            V a = (cl >= 1) ? cb[-S] : V(0);
            V b = (cr <= nc) ? cb[0] : V(0);
            add = V(0.5) * (a + b);
        } else {
            add = cb[0];
        }
        cur[p] += add;
    }
}

// I_fd(prev) at fine point p (prolongation1D_inplace, MatrixVectorHomogen.h:444-474): the
// coarse value below an even fine position, the mean of the two coarse neighbours of an odd
// one (0 outside the coarse box of nc = (nf - 1) / 2 points along the dimension of stride
// S).  Branch-free (neighbouring lanes alternate between the two cases): both reads come from
// in-box addresses and are selected afterwards; the same arithmetic as prolong_dim.
template <typename V>
__device__ inline V prol_add(const V* __restrict__ prev, int p, int S, float rS, int nf, float rnf) {
    const int nc = (nf - 1) >> 1;
    const Split sp = split(p, S, rS, nf, rnf);
    const int j = sp.k >> 1;                    // coarse position below (k - 1) / 2, k = sp.k + 1
    const bool odd = (sp.k & 1) == 0;
    const int ob = sp.lo + mul24(sp.hi, S * nc) + mul24(j, S);
    const bool hasA = odd && j >= 1, hasB = !odd || j < nc;
    const V ra = prev[hasA ? ob - S : 0];
    const V rb = prev[hasB ? ob : 0];
    const V va = hasA ? ra : V(0), vb = hasB ? rb : V(0);
    return odd ? V(0.5) * (va + vb) : vb;
}

// Weighted 1D Q1 operators at level t on one line (positions 1..n, cell c between
// positions c and c+1 carries weight w[c]):
//   (K v)_k = 2^t  [ w_{k-1}(v_k - v_{k-1}) + w_k (v_k - v_{k+1}) ]
//   (M v)_k = 2^-t/6 [ w_{k-1}(v_{k-1} + 2 v_k) + w_k (2 v_k + v_{k+1}) ]
// For a separable coefficient a = prod_j g_j(x_j) the element matrices of
// Poisson::integration (Stencil.h:1707-1789) times a(cell centre) sum to
// sum_i K_i (x) M_{j != i}, applied by the Horner scheme
//   Z <- M_i Z + K_i Y,  Y <- M_i Y   for i = d-1 .. 0.
template <typename V, int PPT, typename SH>
__device__ inline void horner_step(V* __restrict__ Y, V* __restrict__ Z, const SH& sh, int i,
                                   const V* __restrict__ w, int t, bool zeroZ, bool needY) {
    const int S = sh.S[i], n = sh.n[i], total = sh.total;
    const float rS = sh.rS[i], rn = sh.rn[i];
    const V ih = V(1 << t);
    const V h6 = inv6pow2<V>(t);
    V ny[PPT], nz[PPT];
#pragma unroll
    for (int s = 0; s < PPT; ++s) {
        const int p = threadIdx.x + s * blockDim.x;
        if (p < total) {
            const int q = fdiv(p, rS);
            const int kq = q - mul24(fdiv(q, rn), n);
            const bool hasL = kq > 0, hasR = kq < n - 1;
            const V yc = Y[p];
            const V yl = hasL ? Y[p - S] : V(0);
            const V yr = hasR ? Y[p + S] : V(0);
            const V wl = w[kq], wr = w[kq + 1];
            ny[s] = h6 * (wl * (yl + V(2) * yc) + wr * (V(2) * yc + yr));
            V out = ih * (wl * (yc - yl) + wr * (yc - yr));
            if (!zeroZ) {
                const V zc = Z[p];
                const V zl = hasL ? Z[p - S] : V(0);
                const V zr = hasR ? Z[p + S] : V(0);
                out += h6 * (wl * (zl + V(2) * zc) + wr * (V(2) * zc + zr));
            }
            nz[s] = out;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PPT; ++s) {
        const int p = threadIdx.x + s * blockDim.x;
        if (p < total) {
            if (needY) Y[p] = ny[s];
            Z[p] = nz[s];
        }
    }
    __syncthreads();
}

template <typename V, int PPT, typename SH>
__device__ inline void mass_step(V* __restrict__ Y, const SH& sh, int i, const V* __restrict__ w, int t) {
    const int S = sh.S[i], n = sh.n[i], total = sh.total;
    const float rS = sh.rS[i], rn = sh.rn[i];
    const V h6 = inv6pow2<V>(t);
    V ny[PPT];
#pragma unroll
    for (int s = 0; s < PPT; ++s) {
        const int p = threadIdx.x + s * blockDim.x;
        if (p < total) {
            const int q = fdiv(p, rS);
            const int kq = q - mul24(fdiv(q, rn), n);
            const V yc = Y[p];
            const V yl = kq > 0 ? Y[p - S] : V(0);
            const V yr = kq < n - 1 ? Y[p + S] : V(0);
            ny[s] = h6 * (w[kq] * (yl + V(2) * yc) + w[kq + 1] * (V(2) * yc + yr));
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PPT; ++s) {
        const int p = threadIdx.x + s * blockDim.x;
        if (p < total) Y[p] = ny[s];
    }
    __syncthreads();
}

// -------------------------------------------------------------- calcNodal
template <typename V>
__global__ void __launch_bounds__(128) nodal_kernel(const NodalItem* __restrict__ items, int d,
                                                    const V* __restrict__ x, int64_t xStride,
                                                    V* __restrict__ nodal, int64_t nStride, int64_t maxBox,
                                                    const CGState* __restrict__ st, const int* __restrict__ map) {
    extern __shared__ char smem_raw[];
    V* A = reinterpret_cast<V*>(smem_raw);
    V* B = A + maxBox;
    const int inst = map ? map[blockIdx.y] : (int)blockIdx.y;
    if (st && !st[inst].active) return;
    const LSG_CONST NodalItem& it = cptr(items)[blockIdx.x];
    const V* xin = x + inst * xStride + it.in_off;
    V* out = nodal + inst * nStride + it.out_off;
    int cnt = 1, lastE = -1;
    for (int i = 0; i < d; ++i) {
        cnt *= 1 << (it.T[i] - 1);
        if (it.T[i] > 1) lastE = i;
    }
    // along a dimension of level 1 the 1D map is the identity (one coefficient, centre
    // weight 1, no neighbours), so only the other dimensions get a pass; the last pass
    // writes the nodal box
    if (lastE < 0) {
        if (threadIdx.x == 0) out[0] = xin[0];
        return;
    }
    copy_in(A, xin, cnt);
    __syncthreads();
    V* src = A;
    V* dst = B;
    int Slo = 1;
    for (int i = 0; i <= lastE; ++i) {
        const int t = it.T[i];
        if (t == 1) continue;
        const int m = 1 << (t - 1), n = (1 << t) - 1;
        const int outCnt = cnt / m * n;
        expand_dim<V>(src, i == lastE ? out : dst, Slo, m, n, outCnt, t);
        if (i == lastE) break;
        cnt = outCnt;
        Slo *= n;
        __syncthreads();
        V* tmp = src; src = dst; dst = tmp;
    }
}

// ---------------------------------------------------------- prolongation
// P[T+e_dd] += I_dd P[T] along one family (prolongation1D_inplace +
// addMultiLevelValues, MatrixVectorHomogen.h:444-474, 997-1019).
template <typename V>
__global__ void __launch_bounds__(256) prol_kernel(const ProlItem* __restrict__ items, int d,
                                                   const V* __restrict__ nodal, int64_t nStride,
                                                   V* __restrict__ cbuf, int64_t cStride, int64_t maxBox,
                                                   const CGState* __restrict__ st, const int* __restrict__ map) {
    extern __shared__ char smem_raw[];
    V* prev = reinterpret_cast<V*>(smem_raw);
    V* cur = prev + maxBox;
    __shared__ Shape sh;
    const int inst = map ? map[blockIdx.y] : (int)blockIdx.y;
    if (st && !st[inst].active) return;
    const LSG_CONST ProlItem& it = cptr(items)[blockIdx.x];
    const int dd = it.dd;
    const V* nodalI = nodal + inst * nStride;
    const V* cbufI = cbuf + inst * cStride;
    V* outBase = cbuf + inst * cStride;
    int32_t T[MAXD];
    for (int i = 0; i < MAXD; ++i) T[i] = it.T[i];
    for (int q = 0; q < it.nt; ++q) {
        T[dd] = it.T[dd] + q;
        shape_box(sh, T, d);
        const int cnt = sh.total;
        copy_in(cur, (((it.from_nodal >> q) & 1) ? nodalI : cbufI) + it.in_off[q], cnt);
        if (q > 0) {
            __syncthreads();
            prolong_dim<V>(cur, prev, sh, dd);
        }
        __syncthreads();
        if (q > 0) {
            V* out = outBase + it.out_off[q];
            for (int p = threadIdx.x; p < cnt; p += blockDim.x) out[p] = cur[p];
        }
        V* tmp = prev; prev = cur; cur = tmp;
        __syncthreads();
    }
}

// Prolongation chain along dd with the inputs of all its boxes loaded up front: the
// chain's values are dealt out to the threads in chain order (thread t holds chain
// elements t, t + B, t + 2B, ...), so one global round trip covers the whole chain.
// Then box q = input(q) + I_dd(box q-1) (prolongation1D_inplace + addMultiLevelValues,
// MatrixVectorHomogen.h:444-474,997-1019) box by box through LDS, one barrier per box;
// every box but box 0 (whose values stay where they are read from) is written to the
// case buffer.  SC = chain elements per thread; the host guarantees SC * B >= the
// longest chain of the launch.  Grid: item-major, instance fastest.
template <typename V, int SC>
__global__ void __launch_bounds__(512) prol_chain_kernel(const ProlItem* __restrict__ items, int d,
                                                         const V* __restrict__ nodal, int64_t nStride,
                                                         V* __restrict__ cbuf, int64_t cStride, int64_t maxBox,
                                                         int batch, const CGState* __restrict__ st,
                                                         const int* __restrict__ map) {
    extern __shared__ char smem_raw[];
    V* prev = reinterpret_cast<V*>(smem_raw);
    V* cur = prev + maxBox;
    const int inst = map ? map[blockIdx.x % batch] : (int)(blockIdx.x % batch);
    if (st && !st[inst].active) return;
    const LSG_CONST ProlItem& it = cptr(items)[blockIdx.x / batch];
    const int dd = it.dd, nt = it.nt;
    const V* nodalI = nodal + inst * nStride;
    const V* cbufI = cbuf + inst * cStride;
    V* outBase = cbuf + inst * cStride;
    int total0 = 1, S = 1, nf0 = 1;
    for (int i = 0; i < d; ++i) {
        const int n = (1 << it.T[i]) - 1;
        if (i < dd) S *= n;
        if (i == dd) nf0 = n;
        total0 *= n;
    }
    const int B = blockDim.x;
    V nin[SC];
    {
        int off = 0, tq = total0, nq = nf0;
        for (int q = 0; q < nt; ++q) {
            const V* in = (((it.from_nodal >> q) & 1) ? nodalI : cbufI) + it.in_off[q];
#pragma unroll
            for (int s = 0; s < SC; ++s) {
                const int p = threadIdx.x + s * B - off;
                if (p >= 0 && p < tq) nin[s] = in[p];
            }
            off += tq;
            tq = tq / nq * (2 * nq + 1);
            nq = 2 * nq + 1;
        }
    }
    const float rS = rcpf(S);
    int off = 0, tq = total0, nq = nf0;
    for (int q = 0; q < nt; ++q) {
        const bool keep = q + 1 < nt;
        if (q == 0) {
#pragma unroll
            for (int s = 0; s < SC; ++s) {
                const int p = threadIdx.x + s * B;
                if (p < tq) cur[p] = nin[s];
            }
        } else {
            V* out = outBase + it.out_off[q];
            const float rnf = rcpf(nq);
#pragma unroll
            for (int s = 0; s < SC; ++s) {
                const int p = threadIdx.x + s * B - off;
                if (p >= 0 && p < tq) {         // the slots of this box only
                    const V v = nin[s] + prol_add<V>(prev, p, S, rS, nq, rnf);
                    out[p] = v;
                    if (keep) cur[p] = v;
                }
            }
        }
        if (!keep) break;
        __syncthreads();
        V* tmp = prev; prev = cur; cur = tmp;
        off += tq;
        tq = tq / nq * (2 * nq + 1);
        nq = 2 * nq + 1;
    }
}

// Prolongation chain along dd with the input of the next box prefetched into
// registers: box q = input(q) + I_dd(box q-1) (prolongation1D_inplace +
// addMultiLevelValues, MatrixVectorHomogen.h:444-474,997-1019), written to the case
// buffer and kept in LDS for the next box; one barrier per box.  Grid: item-major,
// instance fastest.
template <typename V, int PPT>
__global__ void __launch_bounds__(512) prol_pf_kernel(const ProlItem* __restrict__ items, int d,
                                                      const V* __restrict__ nodal, int64_t nStride,
                                                      V* __restrict__ cbuf, int64_t cStride, int64_t maxBox,
                                                      int batch, const CGState* __restrict__ st,
                                                      const int* __restrict__ map) {
    extern __shared__ char smem_raw[];
    V* prev = reinterpret_cast<V*>(smem_raw);
    V* cur = prev + maxBox;
    const int inst = map ? map[blockIdx.x % batch] : (int)(blockIdx.x % batch);
    if (st && !st[inst].active) return;
    const LSG_CONST ProlItem& it = cptr(items)[blockIdx.x / batch];
    const int dd = it.dd;
    // box q reads the nodal buffer if bit q of from_nodal is set, else the case buffer;
    // box 0 is not written (its values stay where they are read from)
    const V* nodalI = nodal + inst * nStride;
    const V* cbufI = cbuf + inst * cStride;
    V* outBase = cbuf + inst * cStride;
    int total = 1, S = 1, nf = 1;
    for (int i = 0; i < d; ++i) {
        const int n = (1 << it.T[i]) - 1;
        if (i < dd) S *= n;
        if (i == dd) nf = n;
        total *= n;
    }
    V nin[PPT];
    {
        const V* in = ((it.from_nodal & 1) ? nodalI : cbufI) + it.in_off[0];
#pragma unroll
        for (int s = 0; s < PPT; ++s) {
            const int p = threadIdx.x + s * blockDim.x;
            nin[s] = p < total ? in[p] : V(0);
        }
    }
    const float rS = rcpf(S);
    for (int q = 0; q < it.nt; ++q) {
        const bool keep = q + 1 < it.nt;
        V v[PPT];
        if (q == 0) {
#pragma unroll
            for (int s = 0; s < PPT; ++s) v[s] = nin[s];
        } else {
            const float rnf = rcpf(nf);
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                V vs = V(0);
                if (p < total) vs = nin[s] + prol_add<V>(prev, p, S, rS, nf, rnf);
                v[s] = vs;
            }
        }
        const int totalN = keep ? total / nf * (2 * nf + 1) : 0;
        if (keep) {
            const V* in = (((it.from_nodal >> (q + 1)) & 1) ? nodalI : cbufI) + it.in_off[q + 1];
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                nin[s] = p < totalN ? in[p] : V(0);
            }
        }
        V* out = outBase + it.out_off[q];
#pragma unroll
        for (int s = 0; s < PPT; ++s) {
            const int p = threadIdx.x + s * blockDim.x;
            if (p < total) {
                if (q > 0) out[p] = v[s];
                if (keep) cur[p] = v[s];
            }
        }
        if (!keep) break;
        __syncthreads();
        V* tmp = prev; prev = cur; cur = tmp;
        total = totalN;
        nf = 2 * nf + 1;
    }
}

// ------------------------------------------------------------------ apply
// weights of one separable factor for the box's depth: dim i, level T_i,
// 2^T_i cell-centre values, copied from the instance's 1D tables into LDS.
template <typename V, typename TP>
__device__ inline void load_weights(V* __restrict__ W, int* woff, const V* __restrict__ tab, int64_t tabLen,
                                    TP T, int d) {
    int o = 0;
    for (int i = 0; i < d; ++i) {
        if (threadIdx.x == 0) woff[i] = o;
        const V* src = tab + i * tabLen + ((1 << T[i]) - 2);
        for (int c = threadIdx.x; c < (1 << T[i]); c += blockDim.x) W[o + c] = src[c];
        o += 1 << T[i];
    }
    __syncthreads();
}

// One family (T, T+e_fd, ...) of one case: for each box, P = input (+ I_fd(previous
// P)), then z = A_T P, the gM exchanges along the restriction dims and the
// prewavelet conversion into the case-partial output (caseFunction,
// MatrixVectorHomogen.h:965-1090).  LDS: U (P of the previous box), Y, Z
// [, Zt for a sum of terms], weights; each of the box arrays holds a.maxBox values.
template <typename V, int PPT>
__global__ void __launch_bounds__(512) apply_kernel(ApplyArgs<V> a) {
    extern __shared__ char smem_raw[];
    V* U = reinterpret_cast<V*>(smem_raw);
    V* Y = U + a.maxBox;
    V* Z = Y + a.maxBox;
    V* Zt = Z + a.maxBox;                     // general path only
    V* Wt = a.general ? Zt + a.maxBox : Zt;   // weights
    __shared__ int woff[MAXD];
    // family-major order: all instances of a family are adjacent, and the host lists
    // the families of a launch by decreasing work, so the longest start first
    const int inst = a.map ? a.map[blockIdx.x % a.batch] : (int)(blockIdx.x % a.batch);
    if (a.st && !a.st[inst].active) return;
    const LSG_CONST FamilyItem& fam = cptr(a.fam)[blockIdx.x / a.batch];
    const int d = a.d;
    const V* coef = a.coefTab + inst * a.coefStride;
    V* wbase = a.wbuf + inst * a.wStride;

    for (int q = 0; q < fam.nt; ++q) {
        const LSG_CONST ApplyItem& it = cptr(a.items)[fam.first + q];
        // shapes of the box and after each restriction (host-computed, scalar loads)
        const LSG_CONST BoxShape* shs = cptr(a.shapes) + it.shp;
        const LSG_CONST BoxShape& sh = shs[0];
        const V* in = (it.in_nodal ? a.nodal + inst * a.nStride : a.cbuf + inst * a.cStride) + it.in_off;
        copy_in(Y, in, sh.total);
        if (q > 0) {
            __syncthreads();
            prolong_dim<V>(Y, U, sh, fam.fd);
        }
        __syncthreads();
        if (q + 1 < fam.nt || a.general) {        // keep P for the next box / the next term
            copy_in(U, Y, sh.total);
            __syncthreads();
        }
        // stencil: Z = A_T P
        if (!a.general) {
            load_weights<V>(Wt, woff, coef, a.tabLen, it.T, d);
            for (int i = d - 1; i >= 0; --i)
                horner_step<V, PPT>(Y, Z, sh, i, Wt + woff[i], it.T[i], i == d - 1, i > 0);
        } else {
            for (int p = threadIdx.x; p < sh.total; p += blockDim.x) Z[p] = V(0);
            for (int m = 0; m < a.nA + a.nC; ++m) {
                load_weights<V>(Wt, woff, coef + (int64_t)m * d * a.tabLen, a.tabLen, it.T, d);
                if (m > 0) {
                    copy_in(Y, U, sh.total);
                    __syncthreads();
                }
                if (m < a.nA) {
                    for (int i = d - 1; i >= 0; --i)
                        horner_step<V, PPT>(Y, Zt, sh, i, Wt + woff[i], it.T[i], i == d - 1, i > 0);
                    for (int p = threadIdx.x; p < sh.total; p += blockDim.x) Z[p] += Zt[p];
                } else {
                    for (int i = d - 1; i >= 0; --i) mass_step<V, PPT>(Y, sh, i, Wt + woff[i], it.T[i]);
                    for (int p = threadIdx.x; p < sh.total; p += blockDim.x) Z[p] += Y[p];
                }
                __syncthreads();
            }
        }
        // gM exchanges (restriction dims ascending)
        V* cur = Z;
        V* oth = Y;
        if (it.rd_off[0] >= 0) {
            const V* g = wbase + it.rd_off[0];
            for (int p = threadIdx.x; p < sh.total; p += blockDim.x) cur[p] += g[p];
            __syncthreads();
        }
        for (int j = 1; j <= it.k; ++j) {
            const int r = it.r[j - 1];
            restrict_dim<V>(cur, oth, shs[j - 1], r);
            __syncthreads();
            V* tmp = cur; cur = oth; oth = tmp;
            const bool w = it.w_off[j] >= 0, rd = it.rd_off[j] >= 0;
            if (w || rd) {
                V* wd = w ? wbase + it.w_off[j] : nullptr;
                const V* g = rd ? wbase + it.rd_off[j] : nullptr;
                for (int p = threadIdx.x; p < shs[j].total; p += blockDim.x) {
                    V v = cur[p];
                    if (w) wd[p] = v;
                    if (rd) cur[p] = v + g[p];
                }
                __syncthreads();
            }
        }
        // prewavelet conversion, d contractions box -> block
        const LSG_CONST BoxShape& sf = shs[it.k];
        int cnt = sf.total;
        int Slo = 1;
        for (int i = 0; i < d; ++i) {
            const int n = sf.n[i];
            int t = it.T[i];
            for (int j = 0; j < it.k; ++j) if (it.r[j] == i) t -= 1;
            const int m = (n + 1) >> 1;
            const int outCnt = cnt / n * m;
            contract_dim<V>(cur, oth, Slo, n, outCnt, t);
            cnt = outCnt;
            Slo *= m;
            __syncthreads();
            V* tmp = cur; cur = oth; oth = tmp;
        }
        V* out = a.part + inst * a.pStride + it.out_off;
        for (int p = threadIdx.x; p < cnt; p += blockDim.x) out[p] = cur[p];
        __syncthreads();
    }
}

// ------------------------------------------------ apply, one separable term
// Horner sweep of the separable apply kernel.  The first sweep reads P from Y and starts
// Z from zero (P of the box stays in registers for the next box's prolongation, parked by
// the kernel); the last sweep does not store Y, folds in the dimensions of extent 1
// (Z = cM Z + cK Y, see apply_sep_kernel) and adds the gM value g0 read for this box
// (held in registers).
// DEC: bit 0 -- dimension i is the fastest of the box (S = 1, no division for the lower
// part); bit 1 -- no dimension above i has more than one point (S n = total, no division for
// the upper part).  The callers pick it per sweep, so the coordinate decode costs one
// division instead of two on such sweeps.
template <typename V, int PPT, int DEC>
__device__ inline void sep_sweep(V* __restrict__ Y, V* __restrict__ Z, const LSG_CONST BoxShape& sh, int i,
                                 const V* __restrict__ w, int t, bool first, bool needY, bool last, bool fold,
                                 V cM, V cK, const V* __restrict__ g0) {
    const int S = sh.S[i], n = sh.n[i], total = sh.total;
    const float rS = sh.rS[i], rn = sh.rn[i];
    const V ih = V(1 << t);
    const V h6 = inv6pow2<V>(t);
    V ny[PPT], nz[PPT];
    const int B = blockDim.x;
#pragma unroll
    for (int s = 0; s < PPT; ++s) {
        const int p = threadIdx.x + s * B;
        if (p < total) {
            const int q = (DEC & 1) ? p : fdiv(p, rS);
            const int kq = (DEC & 2) ? q : q - mul24(fdiv(q, rn), n);
            const bool hasL = kq > 0, hasR = kq < n - 1;
            // neighbours loaded from an in-box address, then selected
            const V yc = Y[p];
            const V yl0 = Y[hasL ? p - S : p], yr0 = Y[hasR ? p + S : p];
            const V yl = hasL ? yl0 : V(0);
            const V yr = hasR ? yr0 : V(0);
            const V wl = w[kq], wr = w[kq + 1];
            ny[s] = h6 * (wl * (yl + V(2) * yc) + wr * (V(2) * yc + yr));
            V out = ih * (wl * (yc - yl) + wr * (yc - yr));
            if (!first) {
                const V zc = Z[p];
                const V zl0 = Z[hasL ? p - S : p], zr0 = Z[hasR ? p + S : p];
                const V zl = hasL ? zl0 : V(0);
                const V zr = hasR ? zr0 : V(0);
                out += h6 * (wl * (zl + V(2) * zc) + wr * (V(2) * zc + zr));
            }
            nz[s] = out;
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < PPT; ++s) {
        const int p = threadIdx.x + s * B;
        if (p < total) {
            V z = nz[s];
            if (last && fold) {
                const V kz = cK * ny[s];
                z = fma(cM, z, kz);     // explicit, so every instantiation rounds alike
            }
            if (needY && !last) Y[p] = ny[s];
            Z[p] = (last && g0) ? z + g0[p] : z;
        }
    }
    __syncthreads();
}

// Apply stage for an operator with one separable stiffness term (a = prod_j g_j(x_j),
// c = 0): same data flow as apply_kernel, with
//  * the weights of the whole family loaded once (they differ between the boxes of a
//    family only along fd, whose levels T_fd..T_fd+nt-1 are contiguous in the table),
//  * the next box's input and this box's gM value prefetched into registers, so their
//    global latency overlaps the sweeps,
//  * the copy of P, the gM add, the gM write/read of each restriction and the output
//    write fused into the neighbouring passes (one barrier per restriction and
//    contraction, two per Horner sweep, one at the start and one at the end of a box).
template <typename V, int PPT>
__global__ void __launch_bounds__(512) apply_sep_kernel(ApplyArgs<V> a) {
    extern __shared__ char smem_raw[];
    // two LDS boxes A, B and the weights.  At the start of a box, U (one of A, B) holds
    // P of the previous box; P of this box is built in the other one (Y), the Horner
    // sweeps run in place on Y and on U (Z, free once P is built), and at the end of
    // the box P (kept in registers) is parked in the buffer the last pass left free.
    V* U = reinterpret_cast<V*>(smem_raw);
    V* Y = U + a.maxBox;
    V* Z = U;
    V* W = Y + a.maxBox;
    __shared__ int woff[MAXD];
    // family-major order: all instances of a family are adjacent, and the host lists
    // the families of a launch by decreasing work, so the longest start first
    const int inst = a.map ? a.map[blockIdx.x % a.batch] : (int)(blockIdx.x % a.batch);
    if (a.st && !a.st[inst].active) return;
    const LSG_CONST FamilyItem& fam = cptr(a.fam)[blockIdx.x / a.batch];
    const int d = a.d;
    const int fd = fam.fd;
    const V* coef = a.coefTab + inst * a.coefStride;
    const V* nodal = a.nodal + inst * a.nStride;
    const V* cbuf = a.cbuf + inst * a.cStride;
    V* wb = a.wbuf + inst * a.wStride;
    const LSG_CONST ApplyItem& it0 = cptr(a.items)[fam.first];
    const int T0fd = fd >= 0 ? it0.T[fd] : 0;

    // weights of the whole family: dim i at woff[i]; along fd, level T0fd + q at
    // woff[fd] + 2^(T0fd+q) - 2^T0fd
    {
        int o = 0;
        for (int i = 0; i < d; ++i) {
            const int ti = it0.T[i];
            const int len = (i == fd) ? (1 << (ti + fam.nt)) - (1 << ti) : (1 << ti);
            const V* src = coef + i * a.tabLen + ((1 << ti) - 2);
            for (int c = threadIdx.x; c < len; c += blockDim.x) W[o + c] = src[c];
            if (threadIdx.x == 0) woff[i] = o;
            o += len;
        }
    }
    // input of the first box.  FP32: the next box's input is prefetched into registers;
    // FP64: read at the start of each box (the registers cost more occupancy than the
    // prefetch hides, measured)
    constexpr bool kPf = sizeof(V) == 4;
    V nin[kPf ? PPT : 1];
    if constexpr (kPf) {
        int total = 1;
        for (int i = 0; i < d; ++i) total *= (1 << it0.T[i]) - 1;
        const V* in = (it0.in_nodal ? nodal : cbuf) + it0.in_off;
#pragma unroll
        for (int s = 0; s < PPT; ++s) {
            const int p = threadIdx.x + s * blockDim.x;
            nin[s] = p < total ? in[p] : V(0);
        }
    }

    for (int q = 0; q < fam.nt; ++q) {
        const LSG_CONST ApplyItem& it = cptr(a.items)[fam.first + q];
        int total = 1, Sfd = 1, nfd = 1;
        for (int i = 0; i < d; ++i) {
            const int n = (1 << it.T[i]) - 1;
            if (i < fd) Sfd *= n;
            if (i == fd) nfd = n;
            total *= n;
        }
        // P = input (+ I_fd(P of the previous box), prolongation1D_inplace +
        // addMultiLevelValues, MatrixVectorHomogen.h:444-474,997-1019)
        V pk[PPT];      // P of this box, parked for the next box at the end
        const V* in = (it.in_nodal ? nodal : cbuf) + it.in_off;
        // every element of pk is assigned on every path (lanes past the box hold 0), so the
        // register array is not copied around the bounds test
        if (q == 0) {
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                const bool inb = p < total;
                const V v = kPf ? nin[kPf ? s : 0] : (inb ? in[p] : V(0));
                if (inb) Y[p] = v;
                pk[s] = v;
            }
        } else {
            const float rS = rcpf(Sfd), rnf = rcpf(nfd);
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                V v = V(0);
                if (p < total) {
                    v = (kPf ? nin[kPf ? s : 0] : in[p]) + prol_add<V>(U, p, Sfd, rS, nfd, rnf);
                    Y[p] = v;
                }
                pk[s] = v;
            }
        }
        Z = U;
        // gM value of this box (added in the last pass of the stencil); prefetch the
        // input of the next box
        const bool rd0 = it.rd_off[0] >= 0;
        const V* g0 = rd0 ? wb + it.rd_off[0] : nullptr;
        if (kPf && q + 1 < fam.nt) {
            const LSG_CONST ApplyItem& itn = cptr(a.items)[fam.first + q + 1];
            const int totalN = total / nfd * ((nfd << 1) + 1);
            const V* inN = (itn.in_nodal ? nodal : cbuf) + itn.in_off;
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                nin[kPf ? s : 0] = p < totalN ? inN[p] : V(0);
            }
        }
        // shapes of the box and after each restriction (host-computed, scalar loads)
        const LSG_CONST BoxShape* shs = cptr(a.shapes) + it.shp;
        __syncthreads();
        // z = A_T P (Horner sum factorisation), + gM read.  Along a dimension of extent 1
        // (T_i = 1) the 1D operators are the scalars m_i = (w0 + w1)/6 and
        // k_i = 2 (w0 + w1); Horner steps commute, so those dimensions are applied last
        // in closed form: Z = cM A' P + cK M' P with A', M' the stiffness and mass
        // operators of the other dimensions and (cM, cK) from the scalar recurrence.
        const LSG_CONST BoxShape& sh = shs[0];
        const int wq = (1 << (T0fd + q)) - (1 << T0fd);
        int hi = -1, lo = -1;
        bool fold = false;
        V cM = V(1), cK = V(0);
        for (int i = d - 1; i >= 0; --i) {
            if (it.T[i] > 1) {
                if (hi < 0) hi = i;
                lo = i;
            } else {
                fold = true;
                const V* w = W + woff[i] + (i == fd ? wq : 0);
                const V s01 = w[0] + w[1];
                const V m = s01 / V(6), k = V(2) * s01;
                const V kM = k * cM;
                cK = fma(m, cK, kM);
                cM = m * cM;
            }
        }
        const bool keepP = q + 1 < fam.nt;
        if (hi < 0) {
            // every extent is 1 (a single point): Z = cK P
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                if (p < total) {
                    const V yc = Y[p];
                    const V z = cK * yc;
                    Z[p] = rd0 ? z + g0[p] : z;
                }
            }
            __syncthreads();
        } else {
            for (int i = hi; i >= lo; --i) {
                if (it.T[i] == 1) continue;
                const int dec = (sh.S[i] == 1 ? 1 : 0) | (sh.S[i] * sh.n[i] == sh.total ? 2 : 0);
                const V* wi = W + woff[i] + (i == fd ? wq : 0);
                const V* gi = i == lo ? g0 : nullptr;
                switch (dec) {
                    case 0: sep_sweep<V, PPT, 0>(Y, Z, sh, i, wi, it.T[i], i == hi, true, i == lo, fold, cM, cK, gi); break;
                    case 1: sep_sweep<V, PPT, 1>(Y, Z, sh, i, wi, it.T[i], i == hi, true, i == lo, fold, cM, cK, gi); break;
                    case 2: sep_sweep<V, PPT, 2>(Y, Z, sh, i, wi, it.T[i], i == hi, true, i == lo, fold, cM, cK, gi); break;
                    default: sep_sweep<V, PPT, 3>(Y, Z, sh, i, wi, it.T[i], i == hi, true, i == lo, fold, cM, cK, gi); break;
                }
            }
        }
        // restrictions along the restriction dims (ascending) with the gM exchange
        V* cur = Z;
        V* oth = Y;
        for (int j = 1; j <= it.k; ++j) {
            const int r = it.r[j - 1];
            const LSG_CONST BoxShape& s0 = shs[j - 1];
            const int nf = s0.n[r], nc = (nf - 1) >> 1, S = s0.S[r];
            const int outCnt = shs[j].total;
            const float rS = s0.rS[r], rnc = rcpf(nc);
            const bool w = it.w_off[j] >= 0, rd = it.rd_off[j] >= 0;
            V* wd = wb + it.w_off[j];
            const V* g = wb + it.rd_off[j];
            for (int p = threadIdx.x; p < outCnt; p += blockDim.x) {
                Split sp = split(p, S, rS, nc, rnc);
                const V* b = cur + sp.lo + mul24(sp.hi, S * nf) + mul24(2 * sp.k + 1, S);
                const V v = b[0] + V(0.5) * (b[-S] + b[S]);
                if (w) wd[p] = v;
                oth[p] = rd ? v + g[p] : v;
            }
            __syncthreads();
            V* tmp = cur; cur = oth; oth = tmp;
        }
        // prewavelet conversion, d contractions box -> block; the last one writes the part
        // (extent 1 along a dimension means t = 1 there, where the conversion is the identity)
        const LSG_CONST BoxShape& sf = shs[it.k];
        int cnt = sf.total;
        int Slo = 1;
        V* out = a.part + inst * a.pStride + it.out_off;
        int lastC = -1;
        for (int i = 0; i < d; ++i)
            if (sf.n[i] > 1) lastC = i;
        if (lastC < 0) {
            if (threadIdx.x == 0) out[0] = cur[0];
            __syncthreads();
        }
        for (int i = 0; i <= lastC; ++i) {
            const int n = sf.n[i];
            if (n == 1) continue;
            int t = it.T[i];
            for (int j = 0; j < it.k; ++j) if (it.r[j] == i) t -= 1;
            const int m = (n + 1) >> 1;
            const int outCnt = cnt / n * m;
            contract_dim<V>(cur, i < lastC ? oth : out, Slo, n, outCnt, t);
            cnt = outCnt;
            Slo *= m;
            if (i < lastC) {
                __syncthreads();
                V* tmp = cur; cur = oth; oth = tmp;
            }
        }
        // park P in the free buffer (the last pass read cur and wrote global memory)
        if (keepP) {
#pragma unroll
            for (int s = 0; s < PPT; ++s) {
                const int p = threadIdx.x + s * blockDim.x;
                if (p < total) oth[p] = pk[s];
            }
        }
        __syncthreads();
        U = oth;
        Y = cur;
    }
}

// ------------------------------------------------------------------- sum
// Ax = sum over cases of the case partials (MatrixVectorHomogen.h:730-749), in
// ascending case order.
template <typename V>
__global__ void __launch_bounds__(256) sum_kernel(const SumItem* __restrict__ items, const int64_t* __restrict__ list,
                                                  int d, const V* __restrict__ part, int64_t pStride,
                                                  V* __restrict__ y, int64_t yStride,
                                                  const CGState* __restrict__ st, const int* __restrict__ map) {
    const int inst = map ? map[blockIdx.y] : (int)blockIdx.y;
    if (st && !st[inst].active) return;
    const LSG_CONST SumItem& it = cptr(items)[blockIdx.x];
    const int cnt = (int)it.count;
    const V* pb = part + inst * pStride;
    V* out = y + inst * yStride + it.blk_off;
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        V s = V(0);
        for (int c = 0; c < it.ncase; ++c) s = s + pb[cptr(list)[it.list_start + c] + p];
        out[p] = s;
    }
}

// ------------------------------------------------------- preconditioner
// z_p = psi_p^T A_T psi_p for the separable operator (Preconditioning::precond3,
// cg_method.h:322-451): per dimension the 1D quadratic forms of the prewavelet
// against the weighted 1D stiffness and mass matrices.
template <typename V>
__global__ void __launch_bounds__(256) diag_kernel(const SumItem* __restrict__ items, int d, int nA, int nC,
                                                   const V* __restrict__ coefTab, int64_t coefStride, int64_t tabLen,
                                                   V* __restrict__ z, int64_t zStride) {
    const int inst = blockIdx.y;
    const LSG_CONST SumItem& it = cptr(items)[blockIdx.x];
    int m[MAXD];
    int cnt = 1;
    for (int i = 0; i < d; ++i) { m[i] = 1 << (it.T[i] - 1); cnt *= m[i]; }
    const V* coef = coefTab + inst * coefStride;
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        double total = 0.0;
        for (int term = 0; term < nA + nC; ++term) {
            double kq[MAXD], mq[MAXD];
            int rem = p;
            for (int i = 0; i < d; ++i) {
                int j = rem % m[i];
                rem /= m[i];
                const int t = it.T[i];
                const int n = (1 << t) - 1;
                const int k = 2 * j + 1;
                double psi[5];     // positions k-2..k+2
                psi[2] = (t >= 2 && (k == 1 || k == n)) ? 0.9 : 1.0;
                psi[1] = psi[0] = psi[3] = psi[4] = 0.0;
                if (k - 1 >= 1) { psi[1] = -0.6; psi[0] = 0.1; }
                if (k + 1 <= n) { psi[3] = -0.6; psi[4] = 0.1; }
                const V* w = coef + ((int64_t)term * d + i) * tabLen + ((1 << t) - 2);
                double h = 1.0 / double(1 << t);
                double sk = 0.0, sm = 0.0;
                for (int c = k - 3; c <= k + 2; ++c) {      // cell c spans positions c, c+1
                    if (c < 0 || c > n) continue;
                    int a0 = c - (k - 2), a1 = a0 + 1;
                    double va = (a0 >= 0 && a0 <= 4 && c >= 1) ? psi[a0] : 0.0;
                    double vb = (a1 >= 0 && a1 <= 4 && c + 1 <= n) ? psi[a1] : 0.0;
                    double wc = double(w[c]);
                    sk += wc * (vb - va) * (vb - va) / h;
                    sm += wc * h / 3.0 * (va * va + va * vb + vb * vb);
                }
                kq[i] = sk;
                mq[i] = sm;
            }
            if (term < nA) {
                for (int i = 0; i < d; ++i) {
                    double pr = kq[i];
                    for (int j = 0; j < d; ++j) if (j != i) pr *= mq[j];
                    total += pr;
                }
            } else {
                double pr = 1.0;
                for (int j = 0; j < d; ++j) pr *= mq[j];
                total += pr;
            }
        }
        z[inst * zStride + it.blk_off + p] = V(total);
    }
}

// ------------------------------------------------------------------ RHS
// b_p = int f psi_p for f = sum_j eta_j prod_i sin(m_j pi x_i): per dimension
// int sin(w x) hat_q(x) dx = sin(w x_q) 2 (1 - cos(w h)) / (w^2 h).
template <typename V>
__global__ void __launch_bounds__(256) rhs_sines_kernel(const SumItem* __restrict__ items, int d, int nterms,
                                                        const double* __restrict__ eta, const int32_t* __restrict__ mfreq,
                                                        V* __restrict__ b, int64_t bStride,
                                                        const int32_t* __restrict__ pos,
                                                        const uint8_t* __restrict__ active) {
    const int inst = blockIdx.y;
    const LSG_CONST SumItem& it = cptr(items)[blockIdx.x];
    int m[MAXD];
    for (int i = 0; i < d; ++i) m[i] = 1 << (it.T[i] - 1);
    const int cnt = (int)it.count;
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        double total = 0.0;
        for (int term = 0; term < nterms; ++term) {
            const double w = double(mfreq[term]) * M_PI;
            double prod = eta[inst * nterms + term];
            int rem = p;
            for (int i = 0; i < d; ++i) {
                // position of the point at its own level: from the dense block order, or
                // from the position table of a sparse grid
                int j = rem % m[i];
                rem /= m[i];
                const int t = it.T[i];
                const int n = (1 << t) - 1;
                const int k = pos ? pos[(it.blk_off + p) * d + i] : 2 * j + 1;
                const double h = 1.0 / double(1 << t);
                const double fac = 2.0 * (1.0 - cos(w * h)) / (w * w * h);
                double s = ((t >= 2 && (k == 1 || k == n)) ? 0.9 : 1.0) * sin(w * k * h);
                if (k - 1 >= 1) s += -0.6 * sin(w * (k - 1) * h) + 0.1 * sin(w * (k - 2) * h);
                if (k + 1 <= n) s += -0.6 * sin(w * (k + 1) * h) + 0.1 * sin(w * (k + 2) * h);
                prod *= s * fac;
            }
            total += prod;
        }
        // hanging nodes are no unknowns: their entries stay zero (active-only writes of
        // VectorSparseG in the reference)
        if (active && !active[it.blk_off + p]) total = 0.0;
        b[inst * bStride + it.blk_off + p] = V(total);
    }
}

// Load vector of a sum of separable functions: b_p = sum_j w_j prod_i m_ji(T_i, k_i) with
// m_ji(t, k) = int g_ji(x) psi_{t,k}(x) dx the 1D moment of factor (j, i) against the
// prewavelet of odd position k at level t, read from tab[(j d + i) tabLen + 2^(t-1) - 1 +
// (k - 1) / 2] (moments computed on the host).
template <typename V>
__global__ void __launch_bounds__(256) rhs_tab_kernel(const SumItem* __restrict__ items, int d, int nterms,
                                                      const double* __restrict__ w, const double* __restrict__ tab,
                                                      int64_t tabLen, V* __restrict__ b, int64_t bStride,
                                                      const int32_t* __restrict__ pos,
                                                      const uint8_t* __restrict__ active) {
    const int inst = blockIdx.y;
    const LSG_CONST SumItem& it = cptr(items)[blockIdx.x];
    int m[MAXD];
    for (int i = 0; i < d; ++i) m[i] = 1 << (it.T[i] - 1);
    const int cnt = (int)it.count;
    for (int p = threadIdx.x; p < cnt; p += blockDim.x) {
        int64_t off[MAXD];
        int rem = p;
        for (int i = 0; i < d; ++i) {
            const int j = rem % m[i];
            rem /= m[i];
            const int t = it.T[i];
            const int k = pos ? pos[(it.blk_off + p) * d + i] : 2 * j + 1;
            off[i] = (int64_t)(1 << (t - 1)) - 1 + (k - 1) / 2;
        }
        double total = 0.0;
        for (int term = 0; term < nterms; ++term) {
            double prod = w[inst * nterms + term];
            for (int i = 0; i < d; ++i) prod *= tab[(int64_t)(term * d + i) * tabLen + off[i]];
            total += prod;
        }
        if (active && !active[it.blk_off + p]) total = 0.0;   // hanging nodes are no unknowns
        b[inst * bStride + it.blk_off + p] = V(total);
    }
}

// ------------------------------------------------------------- reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
// fixed-order block reduction of up to NV values per thread
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* out) {
    __shared__ double red[NV][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double s = wave_sum(v[q]);
        if (lane == 0) red[q][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += red[q][w];
            out[q] = s;
        }
    }
}

// partials[inst][q][block] of sum_i a_q(i) * b_q(i)
template <typename V, typename W>
__global__ void __launch_bounds__(256) dot_kernel(const V* __restrict__ a, const W* __restrict__ b, int64_t n,
                                                  int64_t stride, double* __restrict__ partials,
                                                  const CGState* __restrict__ st) {
    const int inst = blockIdx.y;
    if (st && !st[inst].active) return;
    double v[1] = {0.0};
    const V* x = a + inst * stride;
    const W* y = b + inst * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[0] += double(x[i]) * double(y[i]);
    double out[1];
    block_sum<1>(v, out);
    if (threadIdx.x == 0) partials[(int64_t)inst * gridDim.x + blockIdx.x] = out[0];
}

__device__ inline double sum_partials(const double* p, int nb) {
    // one thread, fixed order
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += p[b];
    return s;
}

// r = f - r ; h = r / D ; d = -h ; partials: r.h, r.r, f.f.  pre: h, d and r.h are left to the
// depth-block preconditioner (blockpre_kernel)
template <typename V>
__global__ void __launch_bounds__(256) cg_init_kernel(const V* __restrict__ f, V* __restrict__ r, V* __restrict__ h,
                                                      V* __restrict__ dv, const V* __restrict__ D, int64_t n,
                                                      int64_t stride, double* __restrict__ partials, int pre) {
    const int inst = blockIdx.y;
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t o = inst * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        V ri = f[o + i] - r[o + i];
        r[o + i] = ri;
        if (!pre) {
            V hi = ri / D[o + i];
            h[o + i] = hi;
            dv[o + i] = V(-1) * hi;
            v[0] += double(ri) * double(hi);
        }
        v[1] += double(ri) * double(ri);
        v[2] += double(f[o + i]) * double(f[o + i]);
    }
    double out[3];
    block_sum<3>(v, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; ++q) partials[((int64_t)inst * 3 + q) * gridDim.x + blockIdx.x] = out[q];
}

__global__ void cg_start_scalar(CGState* st, const double* __restrict__ partials, int nb, double eps,
                                const double* __restrict__ epsInst, int rule, int maxit,
                                const double* __restrict__ bpart, int nbb) {
    const int inst = blockIdx.x;
    if (threadIdx.x != 0) return;
    CGState& s = st[inst];
    s.delta = bpart ? sum_partials(bpart + (int64_t)inst * nbb, nbb)
                    : sum_partials(partials + ((int64_t)inst * 3 + 0) * nb, nb);
    s.rr = sum_partials(partials + ((int64_t)inst * 3 + 1) * nb, nb);
    s.ff = sum_partials(partials + ((int64_t)inst * 3 + 2) * nb, nb);
    if (epsInst) eps = epsInst[inst];
    s.eps = eps;
    s.rule = rule;
    s.maxit = maxit;
    s.it = 0;
    s.do_dir = 0;
    s.converged = 0;
    if (s.delta != s.delta) { s.active = 0; s.converged = -1; return; }
    bool go = (rule == 0) ? (s.delta > eps) : (s.rr > eps * eps * s.ff);
    s.active = (go && maxit >= 1) ? 1 : 0;
    if (!go) s.converged = 1;
}

__global__ void cg_tau_scalar(CGState* st, const double* __restrict__ partials, int nb) {
    const int inst = blockIdx.x;
    if (threadIdx.x != 0) return;
    CGState& s = st[inst];
    s.do_dir = 0;
    if (!s.active) return;
    double dg = sum_partials(partials + (int64_t)inst * nb, nb);
    s.tau = s.delta / dg;
    s.it += 1;
}

// r += tau g ; x -= tau d ; h = r / D ; partials r.h, r.r (pre: h and r.h by blockpre_kernel)
template <typename V, typename X>
__global__ void __launch_bounds__(256) cg_update_kernel(V* __restrict__ r, X* __restrict__ x, V* __restrict__ h,
                                                        const V* __restrict__ dv, const V* __restrict__ g,
                                                        const V* __restrict__ D, int64_t n, int64_t stride,
                                                        double* __restrict__ partials, const CGState* __restrict__ st,
                                                        int pre) {
    const int inst = blockIdx.y;
    if (!st[inst].active) return;
    const V tau = V(st[inst].tau);
    double v[2] = {0.0, 0.0};
    const int64_t o = inst * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        V ri = r[o + i] + tau * g[o + i];
        x[o + i] = x[o + i] - X(tau) * X(dv[o + i]);
        r[o + i] = ri;
        if (!pre) {
            V hi = V(1) * ri / D[o + i];
            h[o + i] = hi;
            v[0] += double(ri) * double(hi);
        }
        v[1] += double(ri) * double(ri);
    }
    double out[2];
    block_sum<2>(v, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < 2; ++q) partials[((int64_t)inst * 2 + q) * gridDim.x + blockIdx.x] = out[q];
}

__global__ void cg_beta_scalar(CGState* st, const double* __restrict__ partials, int nb,
                               const double* __restrict__ bpart, int nbb) {
    const int inst = blockIdx.x;
    if (threadIdx.x != 0) return;
    CGState& s = st[inst];
    if (!s.active) return;
    double dp = bpart ? sum_partials(bpart + (int64_t)inst * nbb, nbb)
                      : sum_partials(partials + ((int64_t)inst * 2 + 0) * nb, nb);
    s.rr = sum_partials(partials + ((int64_t)inst * 2 + 1) * nb, nb);
    if (dp != dp) { s.active = 0; s.converged = -1; s.delta = dp; return; }
    bool stop = (s.rule == 0) ? (dp < s.eps * s.eps) : (s.rr <= s.eps * s.eps * s.ff);
    if (stop) { s.active = 0; s.converged = 1; s.delta = dp; return; }
    s.beta = dp / s.delta;
    s.delta = dp;
    s.do_dir = 1;
    bool cont = (s.rule == 0) ? (s.delta > s.eps) : true;
    if (!cont || s.it >= s.maxit) {
        s.active = 0;
        s.converged = cont ? 0 : 1;
    }
}

// Depth-block preconditioner (the inner solves of the mixed-precision method; no counterpart
// in the reference, whose CG divides by the diagonal): workgroup (block T, group of ipw
// instances, a multiple of 4); the block's inverse is staged in LDS once and applied by each wave to one
// instance at a time: lane k forms h_k = s_k sum_j B^-1_kj s_j r_j for the points k = lane and
// k = lane + 64 of a block of up to 128 points (x = s .* r broadcast from LDS, the inverse read
// column-wise), s = sqrt(d0 / D).  Blocks without an inverse: h = r / D.  The r.h partial of
// (instance, block) is summed over the wave in a fixed order.  init: also d = -h (the start of
// a solve, every instance).
constexpr int PRE_MAXN = 128;
template <typename V, bool INIT>
__global__ void __launch_bounds__(256) blockpre_kernel(const PreBlock* __restrict__ items, int nblk,
                                                       const V* __restrict__ inv, const V* __restrict__ d0,
                                                       const V* __restrict__ r, const V* __restrict__ D,
                                                       V* __restrict__ h, V* __restrict__ dv, int64_t n, int batch,
                                                       int ipw, double* __restrict__ part,
                                                       const CGState* __restrict__ st, int mmMin) {
    extern __shared__ char smem_raw[];
    const PreBlock it = items[blockIdx.x];
    if (it.mode == PRE_FD) return;         // fast-diagonalisation block: blockfd_kernel
    if (it.mode == PRE_DENSE && it.n >= mmMin) return;     // blockmm_kernel
    const int nl = it.n;
    V* Bs = reinterpret_cast<V*>(smem_raw);
    const bool dense = it.mode == PRE_DENSE;
    V* xs = Bs + (dense ? nl * nl : 0);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (dense) {
        for (int e = threadIdx.x; e < nl * nl; e += blockDim.x) Bs[e] = inv[(int64_t)it.inv + e];
    }
    __syncthreads();
    V* x = xs + wave * PRE_MAXN;
    const int i0 = blockIdx.y * ipw;
    for (int q = 0; q < ipw; q += 4) {
        const int inst = i0 + q + wave;
        const bool run = inst < batch && (INIT || st[inst].active);
        const int64_t o = (int64_t)inst * n + it.off;
        double acc = 0.0;
        if (dense) {
            // nl <= 128: points lane and lane + 64
            V rk[2] = {V(0), V(0)}, sk[2] = {V(0), V(0)};
            if (run) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = lane + 64 * u;
                    if (k < nl) {
                        rk[u] = r[o + k];
                        sk[u] = sqrt(d0[it.off + k] / D[o + k]);
                        x[k] = sk[u] * rk[u];
                    }
                }
            }
            __syncthreads();
            if (run) {
                V a0 = V(0), a1 = V(0);
                const bool two = lane + 64 < nl;
                if (lane < nl) {
                    for (int j = 0; j < nl; ++j) {
                        const V xj = x[j];
                        a0 = fma(Bs[j * nl + lane], xj, a0);
                        if (two) a1 = fma(Bs[j * nl + lane + 64], xj, a1);
                    }
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int k = lane + 64 * u;
                    if (k < nl) {
                        const V hk = sk[u] * (u ? a1 : a0);
                        h[o + k] = hk;
                        if (INIT) dv[o + k] = V(-1) * hk;
                        acc += double(rk[u]) * double(hk);
                    }
                }
            }
            __syncthreads();
        } else if (run) {
            for (int k = lane; k < nl; k += 64) {
                const V rk = r[o + k];
                const V hk = rk / D[o + k];
                h[o + k] = hk;
                if (INIT) dv[o + k] = V(-1) * hk;
                acc += double(rk) * double(hk);
            }
        }
        if (run) {
            acc = wave_sum(acc);
            if (lane == 0) part[(int64_t)inst * nblk + blockIdx.x] = acc;
        }
    }
}

// FP32 dense blocks of at least PRE_MM_MIN points on the matrix cores: Y = B^-1 X for a group of
// 64 instances is a (n x n) x (n x 64) product, formed with v_mfma_f32_32x32x2_f32 (FP32
// operands and accumulation) over 32 x 32 output tiles, the block padded to a multiple of
// 32 with zeros.  Workgroup (block, `groups` groups of 64 instances): X = s .* r is staged in
// LDS (instance-major, stride n + 1, so the 32 instances a wave reads fall in distinct
// banks); the A operand comes straight from the inverse in global memory, which every group
// reads and so stays in L2, leaving the LDS to X alone (33 KB: four workgroups per CU).  Each
// wave takes output tiles (row strip, 32 instances) and accumulates over k in a fixed order,
// so an instance's result does not depend on the others in its group.  The tiles are parked
// in LDS and written out instance by instance (lanes along the block's points), with
// h = s .* y and the r.h partial summed over the wave in a fixed order as in blockpre_kernel.
typedef float floatx16_t __attribute__((ext_vector_type(16)));
template <bool INIT>
__global__ void __launch_bounds__(256) blockmm_kernel(BlockPreArgs<float> a, const float* __restrict__ r,
                                                      const float* __restrict__ D, float* __restrict__ h,
                                                      float* __restrict__ dv, int64_t n, int batch, int groups,
                                                      const CGState* __restrict__ st) {
    extern __shared__ char smem_raw[];
    const int bi = a.mmList[blockIdx.x];
    const PreBlock it = a.items[bi];
    const int nl = it.n, np = (nl + 31) & ~31, ld = np + 1;
    float* X = reinterpret_cast<float*>(smem_raw);     // 64 x ld, instance-major (then Y)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float* inv = a.inv + it.inv;                   // column k contiguous: (i, k) at k * nl + i
    // output tiles: wave w takes (row strip R0 = w / 2, instances C = w % 2) and, for blocks
    // of 128 points, also row strip R0 + 2 with the same instances (one B read, two MFMAs)
    const int nR = np >> 5;
    const int C = wave & 1, R0 = wave >> 1, R1 = R0 + 2;
    const bool t0 = R0 < nR, t1 = R1 < nR;
    for (int grp = 0; grp < groups; ++grp) {
        const int i0 = (blockIdx.y * groups + grp) * 64;
        if (i0 >= batch) break;
        // X = s .* r, four instances per wave and step
        for (int c0 = wave; c0 < 64; c0 += 16) {
            float v[4][2];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int inst = i0 + c0 + 4 * u;
                const bool run = inst < batch && (INIT || st[inst].active);
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    const int k = lane + 64 * w;
                    float x = 0.0f;
                    if (run && k < nl) {
                        const int64_t o = (int64_t)inst * n + it.off + k;
                        x = sqrtf(a.d0[it.off + k] / D[o]) * r[o];
                    }
                    v[u][w] = x;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 2; ++w) {
                    const int k = lane + 64 * w;
                    if (k < np) X[(c0 + 4 * u) * ld + k] = v[u][w];
                }
        }
        __syncthreads();
        // lane holds instance C * 32 + (lane & 31) and rows R * 32 + 8 (q / 4) + 4 (lane / 32)
        // + q % 4 of its 16 accumulators; k in a fixed order
        floatx16_t c0v = {}, c1v = {};
        {
            // A operand straight from the inverse in global memory (L2-resident: every group
            // of instances reads the same block): lanes 0-31 read rows of column k0, lanes
            // 32-63 of column k0 + 1, two 128-byte segments; rows and columns past the block
            // (padding to 32) read an in-block address and take 0
            const int row0 = R0 * 32 + (lane & 31), row1 = R1 * 32 + (lane & 31);
            const bool in0 = row0 < nl, in1 = row1 < nl;
            const float* ar0 = inv + (in0 ? row0 : 0) + (int64_t)(lane >> 5) * nl;
            const float* ar1 = inv + (in1 ? row1 : 0) + (int64_t)(lane >> 5) * nl;
            const float* br = X + (C * 32 + (lane & 31)) * ld + (lane >> 5);
            const int kEnd = nl - (lane >> 5);     // this lane's columns k0 + lane / 32 < nl
            if (t1) {
#pragma unroll 4
                for (int k0 = 0; k0 < np; k0 += 2) {
                    const bool kin = k0 < kEnd;
                    const float x0 = ar0[kin ? (int64_t)k0 * nl : 0], x1 = ar1[kin ? (int64_t)k0 * nl : 0];
                    const float b = br[k0];
                    c0v = __builtin_amdgcn_mfma_f32_32x32x2f32((kin && in0) ? x0 : 0.0f, b, c0v, 0, 0, 0);
                    c1v = __builtin_amdgcn_mfma_f32_32x32x2f32((kin && in1) ? x1 : 0.0f, b, c1v, 0, 0, 0);
                }
            } else if (t0) {
#pragma unroll 4
                for (int k0 = 0; k0 < np; k0 += 2) {
                    const bool kin = k0 < kEnd;
                    const float x0 = ar0[kin ? (int64_t)k0 * nl : 0];
                    c0v = __builtin_amdgcn_mfma_f32_32x32x2f32((kin && in0) ? x0 : 0.0f, br[k0], c0v, 0, 0, 0);
                }
            }
        }
        __syncthreads();
        {
            float* y = X + (C * 32 + (lane & 31)) * ld + 4 * (lane >> 5);
            if (t0) {
#pragma unroll
                for (int q = 0; q < 16; ++q) y[R0 * 32 + 8 * (q >> 2) + (q & 3)] = c0v[q];
            }
            if (t1) {
#pragma unroll
                for (int q = 0; q < 16; ++q) y[R1 * 32 + 8 * (q >> 2) + (q & 3)] = c1v[q];
            }
        }
        __syncthreads();
        for (int c = wave; c < 64; c += 4) {
            const int inst = i0 + c;
            if (!(inst < batch && (INIT || st[inst].active))) continue;
            const int64_t o = (int64_t)inst * n + it.off;
            double accd = 0.0;
            for (int k = lane; k < nl; k += 64) {
                const float rk = r[o + k];
                const float hk = sqrtf(a.d0[it.off + k] / D[o + k]) * X[c * ld + k];
                h[o + k] = hk;
                if (INIT) dv[o + k] = -1.0f * hk;
                accd += double(rk) * double(hk);
            }
            accd = wave_sum(accd);
            if (lane == 0) a.part[(int64_t)inst * a.nblk + bi] = accd;
        }
        __syncthreads();
    }
}

// The larger depth blocks by fast diagonalisation of the Kronecker sum:
// h = s .* (V Lam^-1 V^T (s .* r)) with V = (x)_i V_i over the block's product set of 1D
// positions S_1 x ... x S_d, applied mode by mode in LDS (the product set ordered dimension 0
// fastest, |S_i| along dimension i; a point order table maps it to the block's device order
// when they differ, and lists the hanging points after it, which take h = r / D).  Workgroup
// (block, group of ipw instances); each wave transforms one instance at a time in its own two
// LDS buffers; every wave of a workgroup runs the same passes, so the passes are separated by
// workgroup barriers.
template <typename V, bool INIT>
__global__ void __launch_bounds__(256) blockfd_kernel(BlockPreArgs<V> a, const V* __restrict__ r,
                                                      const V* __restrict__ D, V* __restrict__ h, V* __restrict__ dv,
                                                      int64_t n, int batch, int ipw, const CGState* __restrict__ st) {
    extern __shared__ char smem_raw[];
    const int bi = a.fdList[blockIdx.x];
    const PreBlock it = a.items[bi];
    const int nl = it.n, np = it.np, d = a.d;
    const int32_t* perm = it.inv >= 0 ? a.fdPerm + it.inv : nullptr;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    V* A = reinterpret_cast<V*>(smem_raw) + (int64_t)wave * 2 * a.maxFd;
    V* B = A + a.maxFd;
    int str[MAXD], sh[MAXD];
    bool p2 = true;                        // every extent a power of two: coordinates are bit fields
    V f = V(1);                            // extents of 1: V_i = v (1 x 1) forward and backward
    for (int i = 0, o = 1; i < d; ++i) {
        const int m = it.m[i];
        str[i] = o;
        sh[i] = __builtin_ctz((unsigned)o);
        p2 = p2 && (m & (m - 1)) == 0;
        o *= m;
        if (m == 1) {
            const V v = a.fdV[it.fv[i]];
            f *= v * v;
        }
    }
    auto coord = [&](int p, int i) { return p2 ? (p >> sh[i]) & (it.m[i] - 1) : (p / str[i]) % it.m[i]; };
    const int i0 = blockIdx.y * ipw;
    for (int q = 0; q < ipw; q += 4) {
        const int inst = i0 + q + wave;
        const bool run = inst < batch && (INIT || st[inst].active);
        const int64_t ob = (int64_t)inst * n + it.off;
        V* X = A;
        V* Y = B;
        if (run)
            for (int p = lane; p < np; p += 64) {
                const int64_t k = perm ? perm[p] : p;
                X[p] = sqrt(a.d0[it.off + k] / D[ob + k]) * r[ob + k];
            }
        __syncthreads();
        // forward: along every dimension with more than one position, X <- V_i^T X
        for (int i = 0; i < d; ++i) {
            const int m = it.m[i];
            if (m == 1) continue;
            const V* Vt = a.fdV + it.fv[i];
            if (run)
                for (int p = lane; p < np; p += 64) {
                    const int k = coord(p, i);
                    const int base = p - k * str[i];
                    V acc = V(0);
                    for (int j = 0; j < m; ++j) acc = fma(Vt[j * m + k], X[base + j * str[i]], acc);
                    Y[p] = acc;
                }
            __syncthreads();
            V* t = X; X = Y; Y = t;
        }
        // 1 / (sum of the 1D eigenvalues), times the factors of the extents of 1
        if (run)
            for (int p = lane; p < np; p += 64) {
                V ls = V(0);
                for (int i = 0; i < d; ++i) ls += a.fdLam[it.fl[i] + coord(p, i)];
                X[p] = f * X[p] / ls;
            }
        __syncthreads();
        // backward: X <- V_i X along the same dimensions
        for (int i = 0; i < d; ++i) {
            const int m = it.m[i];
            if (m == 1) continue;
            const V* Vt = a.fdV + it.fv[i];
            if (run)
                for (int p = lane; p < np; p += 64) {
                    const int k = coord(p, i);
                    const int base = p - k * str[i];
                    V acc = V(0);
                    for (int j = 0; j < m; ++j) acc = fma(Vt[k * m + j], X[base + j * str[i]], acc);
                    Y[p] = acc;
                }
            __syncthreads();
            V* t = X; X = Y; Y = t;
        }
        double acc = 0.0;
        if (run) {
            for (int p = lane; p < nl; p += 64) {
                const int64_t k = perm ? perm[p] : p;
                const V rk = r[ob + k];
                const V hk = p < np ? sqrt(a.d0[it.off + k] / D[ob + k]) * X[p] : rk / D[ob + k];
                h[ob + k] = hk;
                if (INIT) dv[ob + k] = V(-1) * hk;
                acc += double(rk) * double(hk);
            }
            acc = wave_sum(acc);
            if (lane == 0) a.part[(int64_t)inst * a.nblk + bi] = acc;
        }
        __syncthreads();
    }
}

template <typename V>
__global__ void __launch_bounds__(256) cg_dir_kernel(V* __restrict__ dv, const V* __restrict__ h, int64_t n,
                                                     int64_t stride, const CGState* __restrict__ st) {
    const int inst = blockIdx.y;
    if (!st[inst].do_dir) return;
    const V beta = V(st[inst].beta);
    const int64_t o = inst * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dv[o + i] = V(-1) * h[o + i] + beta * dv[o + i];
}

// elementwise helpers for iterative refinement
template <typename A, typename B>
__global__ void __launch_bounds__(256) cast_kernel(const A* __restrict__ a, B* __restrict__ b, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = B(a[i]);
}
// r = f - Ax (FP64), partials of r.r and f.f per instance
__global__ void __launch_bounds__(256) residual_kernel(const double* __restrict__ f, const double* __restrict__ Ax,
                                                       double* __restrict__ r, int64_t n, int64_t stride,
                                                       double* __restrict__ partials, const int* __restrict__ map) {
    const int inst = map ? map[blockIdx.y] : (int)blockIdx.y;
    double v[2] = {0.0, 0.0};
    const int64_t o = inst * stride;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double ri = f[o + i] - Ax[o + i];
        r[o + i] = ri;
        v[0] += ri * ri;
        v[1] += f[o + i] * f[o + i];
    }
    double out[2];
    block_sum<2>(v, out);
    if (threadIdx.x == 0)
        for (int q = 0; q < 2; ++q) partials[((int64_t)inst * 2 + q) * gridDim.x + blockIdx.x] = out[q];
}
// x += e (FP32 correction into the FP64 iterate)
__global__ void __launch_bounds__(256) axpy_correction_kernel(double* __restrict__ x, const float* __restrict__ e,
                                                              int64_t total) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x)
        x[i] += double(e[i]);
}
template <typename V>
__global__ void __launch_bounds__(256) fill_one_kernel(V* __restrict__ p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = V(1);
}

__global__ void __launch_bounds__(256) zero_kernel(void* __restrict__ p, int64_t bytes) {
    uint32_t* q = reinterpret_cast<uint32_t*>(p);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < bytes / 4; i += (int64_t)gridDim.x * blockDim.x)
        q[i] = 0u;
}

// rows[j] of a (rows x n) FP64 array set to zero, for j < nrows (grid.y = nrows)
__global__ void __launch_bounds__(256) zero_rows_kernel(double* __restrict__ a, int64_t n,
                                                        const int* __restrict__ rows) {
    double* r = a + rows[blockIdx.y] * n;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = 0.0;
}

// ------------------------------------------------------------ launchers
static inline unsigned gridFor(int64_t n, int per = 256 * 4) {
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > 1024) g = 1024;
    return (unsigned)g;
}

template <typename V>
void launch_nodal(const NodalItem* items, int nitems, int d, const V* x, int64_t xStride, V* nodal, int64_t nStride,
                  int64_t maxBox, int batch, const CGState* st, hipStream_t s, const int* map) {
    if (nitems == 0) return;
    size_t lds = 2 * maxBox * sizeof(V);
    hipLaunchKernelGGL(nodal_kernel<V>, dim3(nitems, batch), dim3(128), lds, s, items, d, x, xStride, nodal, nStride,
                       maxBox, st, map);
}
template <typename V>
void launch_prol(const ProlItem* items, int nitems, int d, const V* nodal, int64_t nStride, V* cbuf, int64_t cStride,
                 int64_t maxBox, int64_t maxChain, int batch, const CGState* st, hipStream_t s, const int* map) {
    if (nitems == 0) return;
    const size_t lds = 2 * maxBox * sizeof(V);
    if (sizeof(V) == 4) {
        // FP32: whole chains in registers, in the narrowest workgroup that holds the
        // longest chain with at most 24 values per thread (FP64 keeps the per-box
        // prefetch below: there the registers cost more occupancy than the single
        // round trip saves, measured)
        int threads = 128;
        while (threads < 512 && maxChain > 24 * (int64_t)threads) threads *= 2;
        const int64_t sc = (maxChain + threads - 1) / threads;
        const dim3 grid((unsigned)nitems * (unsigned)batch), blk(threads);
#define LSG_CHAIN_CASE(P)                                                                                     \
        if (sc <= P) {                                                                                        \
            hipLaunchKernelGGL((prol_chain_kernel<V, P>), grid, blk, lds, s, items, d, nodal, nStride, cbuf,   \
                               cStride, maxBox, batch, st, map);                                              \
            return;                                                                                           \
        }
        LSG_CHAIN_CASE(2)
        LSG_CHAIN_CASE(4)
        LSG_CHAIN_CASE(6)
        LSG_CHAIN_CASE(8)
        LSG_CHAIN_CASE(12)
        LSG_CHAIN_CASE(16)
        LSG_CHAIN_CASE(20)
        LSG_CHAIN_CASE(24)
#undef LSG_CHAIN_CASE
    }
    // the narrowest workgroup whose 16 points per thread cover the largest box (a chain
    // is serial, so more, smaller workgroups hide its latency better; measured)
    const int threads = maxBox <= 16 * 128 ? 128 : (maxBox <= 16 * 256 ? 256 : 512);
    const int64_t ppt = (maxBox + threads - 1) / threads;
    const dim3 grid((unsigned)nitems * (unsigned)batch), blk(threads);
#define LSG_PROL_CASE(P)                                                                                   \
    if (ppt <= P) {                                                                                        \
        hipLaunchKernelGGL((prol_pf_kernel<V, P>), grid, blk, lds, s, items, d, nodal, nStride, cbuf, cStride, \
                           maxBox, batch, st, map);                                                        \
        return;                                                                                            \
    }
    LSG_PROL_CASE(1)
    LSG_PROL_CASE(2)
    LSG_PROL_CASE(3)
    LSG_PROL_CASE(4)
    LSG_PROL_CASE(6)
    LSG_PROL_CASE(8)
    LSG_PROL_CASE(12)
    LSG_PROL_CASE(16)
#undef LSG_PROL_CASE
    hipLaunchKernelGGL(prol_kernel<V>, dim3(nitems, batch), dim3(256), lds, s, items, d, nodal, nStride, cbuf, cStride,
                       maxBox, st, map);
}
template <typename V>
int launch_apply(const ApplyArgs<V>& a, int nitems, int batch, int threads, size_t lds, hipStream_t s) {
    if (nitems == 0) return 0;
    const int64_t ppt = (a.maxBox + threads - 1) / threads;
    dim3 grid((unsigned)nitems * (unsigned)batch), blk(threads);
#define LSG_APPLY_CASE(P)                                                       \
    if (ppt <= P) {                                                             \
        if (a.general)                                                          \
            hipLaunchKernelGGL((apply_kernel<V, P>), grid, blk, lds, s, a);     \
        else                                                                    \
            hipLaunchKernelGGL((apply_sep_kernel<V, P>), grid, blk, lds, s, a); \
        return 0;                                                               \
    }
    LSG_APPLY_CASE(1)
    LSG_APPLY_CASE(2)
    LSG_APPLY_CASE(3)
    LSG_APPLY_CASE(4)
    LSG_APPLY_CASE(5)
    LSG_APPLY_CASE(6)
    LSG_APPLY_CASE(8)
    LSG_APPLY_CASE(12)
    LSG_APPLY_CASE(16)
#undef LSG_APPLY_CASE
    return -1;
}
template <typename V>
void launch_sum(const SumItem* items, int nitems, const int64_t* list, int d, const V* part, int64_t pStride, V* y,
                int64_t yStride, int batch, const CGState* st, hipStream_t s, const int* map) {
    // one wave per depth block: the blocks are small (2^(|T|_1 - d) values, 64 at most
    // for d=5 level 7), so wider workgroups would mostly idle (measured)
    hipLaunchKernelGGL(sum_kernel<V>, dim3(nitems, batch), dim3(64), 0, s, items, list, d, part, pStride, y, yStride,
                       st, map);
}
template <typename V>
void launch_diag(const SumItem* items, int nitems, int d, int nA, int nC, const V* coefTab, int64_t coefStride,
                 int64_t tabLen, V* z, int64_t zStride, int batch, hipStream_t s) {
    hipLaunchKernelGGL(diag_kernel<V>, dim3(nitems, batch), dim3(128), 0, s, items, d, nA, nC, coefTab, coefStride,
                       tabLen, z, zStride);
}
template <typename V>
void launch_rhs_sines(const SumItem* items, int nitems, int d, int nterms, const double* eta, const int32_t* m, V* b,
                      int64_t bStride, int batch, hipStream_t s, const int32_t* pos, const uint8_t* active) {
    hipLaunchKernelGGL(rhs_sines_kernel<V>, dim3(nitems, batch), dim3(128), 0, s, items, d, nterms, eta, m, b, bStride,
                       pos, active);
}

template <typename V>
void launch_rhs_tab(const SumItem* items, int nitems, int d, int nterms, const double* w, const double* tab,
                    int64_t tabLen, V* b, int64_t bStride, int batch, hipStream_t s, const int32_t* pos,
                    const uint8_t* active) {
    hipLaunchKernelGGL(rhs_tab_kernel<V>, dim3(nitems, batch), dim3(128), 0, s, items, d, nterms, w, tab, tabLen, b,
                       bStride, pos, active);
}

template <typename V>
static void launch_blockpre(const BlockPreArgs<V>& pre, bool init, const V* r, const V* D, V* h, V* dv, int64_t n,
                            int batch, const CGState* st, hipStream_t s) {
    // instances per workgroup: the staged inverse (up to 64 KB) serves up to 64 of them
    const int ipw = std::min(64, (batch + 3) / 4 * 4);
    const dim3 grid((unsigned)pre.nblk, (unsigned)((batch + ipw - 1) / ipw));
    const size_t lds = ((size_t)pre.maxN * pre.maxN + 4 * PRE_MAXN) * sizeof(V);
    const int mmMin = pre.nmm > 0 ? PRE_MM_MIN : 1 << 30;
    if (init)
        hipLaunchKernelGGL((blockpre_kernel<V, true>), grid, dim3(256), lds, s, pre.items, pre.nblk, pre.inv, pre.d0, r,
                           D, h, dv, n, batch, ipw, pre.part, st, mmMin);
    else
        hipLaunchKernelGGL((blockpre_kernel<V, false>), grid, dim3(256), lds, s, pre.items, pre.nblk, pre.inv, pre.d0,
                           r, D, h, dv, n, batch, ipw, pre.part, st, mmMin);
    if constexpr (std::is_same<V, float>::value) {
        if (pre.nmm > 0) {
            const int np = (pre.mmMax + 31) & ~31;
            const size_t ldm = (size_t)64 * (np + 1) * sizeof(float);
            // groups of 64 instances per workgroup (LSG_PRE_MM_GROUPS, default 1)
            const int ng = (batch + 63) / 64, groups = std::max(1, std::min(pre.mmGroups, ng));
            const dim3 gmm((unsigned)pre.nmm, (unsigned)((ng + groups - 1) / groups));
            if (init)
                hipLaunchKernelGGL((blockmm_kernel<true>), gmm, dim3(256), ldm, s, pre, r, D, h, dv, n, batch, groups,
                                   st);
            else
                hipLaunchKernelGGL((blockmm_kernel<false>), gmm, dim3(256), ldm, s, pre, r, D, h, dv, n, batch,
                                   groups, st);
        }
    }
    if (pre.nfd > 0) {
        // 16 instances per workgroup: the blocks are few and large
        const int ipf = std::min(16, (batch + 3) / 4 * 4);
        const dim3 gfd((unsigned)pre.nfd, (unsigned)((batch + ipf - 1) / ipf));
        const size_t ldf = (size_t)8 * pre.maxFd * sizeof(V);
        if (init)
            hipLaunchKernelGGL((blockfd_kernel<V, true>), gfd, dim3(256), ldf, s, pre, r, D, h, dv, n, batch, ipf, st);
        else
            hipLaunchKernelGGL((blockfd_kernel<V, false>), gfd, dim3(256), ldf, s, pre, r, D, h, dv, n, batch, ipf, st);
    }
}
template <typename V>
void launch_cg_init(const V* f, V* r, V* h, V* dv, const V* D, int64_t n, int batch, double* partials, int nb,
                    hipStream_t s, const BlockPreArgs<V>& pre) {
    hipLaunchKernelGGL(cg_init_kernel<V>, dim3(nb, batch), dim3(256), 0, s, f, r, h, dv, D, n, n, partials,
                       pre.items ? 1 : 0);
    if (pre.items) launch_blockpre<V>(pre, true, r, D, h, dv, n, batch, nullptr, s);
}
void launch_cg_start(CGState* st, const double* partials, int nb, double eps, const double* epsInst, int rule,
                     int maxit, int batch, hipStream_t s, const double* bpart, int nbb) {
    hipLaunchKernelGGL(cg_start_scalar, dim3(batch), dim3(64), 0, s, st, partials, nb, eps, epsInst, rule, maxit,
                       bpart, nbb);
}
template <typename V>
void launch_cg_iter_tail(V* r, void* x, bool x64, V* h, V* dv, const V* g, const V* D, int64_t n, int batch,
                         double* partials, int nb, CGState* st, hipStream_t s, const BlockPreArgs<V>& pre) {
    const int pf = pre.items ? 1 : 0;
    hipLaunchKernelGGL((dot_kernel<V, V>), dim3(nb, batch), dim3(256), 0, s, dv, g, n, n, partials, st);
    hipLaunchKernelGGL(cg_tau_scalar, dim3(batch), dim3(64), 0, s, st, partials, nb);
    if (x64)
        hipLaunchKernelGGL((cg_update_kernel<V, double>), dim3(nb, batch), dim3(256), 0, s, r, (double*)x, h, dv, g, D,
                           n, n, partials, st, pf);
    else
        hipLaunchKernelGGL((cg_update_kernel<V, V>), dim3(nb, batch), dim3(256), 0, s, r, (V*)x, h, dv, g, D, n, n,
                           partials, st, pf);
    if (pf) launch_blockpre<V>(pre, false, r, D, h, dv, n, batch, st, s);
    hipLaunchKernelGGL(cg_beta_scalar, dim3(batch), dim3(64), 0, s, st, partials, nb, pf ? pre.part : nullptr,
                       pre.nblk);
    hipLaunchKernelGGL(cg_dir_kernel<V>, dim3(nb, batch), dim3(256), 0, s, dv, h, n, n, st);
}
void launch_residual(const double* f, const double* Ax, double* r, int64_t n, int batch, double* partials, int nb,
                     hipStream_t s, const int* map) {
    hipLaunchKernelGGL(residual_kernel, dim3(nb, batch), dim3(256), 0, s, f, Ax, r, n, n, partials, map);
}
void launch_cast_d2f(const double* a, float* b, int64_t total, hipStream_t s) {
    hipLaunchKernelGGL((cast_kernel<double, float>), dim3(gridFor(total)), dim3(256), 0, s, a, b, total);
}
void launch_cast_f2d(const float* a, double* b, int64_t total, hipStream_t s) {
    hipLaunchKernelGGL((cast_kernel<float, double>), dim3(gridFor(total)), dim3(256), 0, s, a, b, total);
}
void launch_axpy_correction(double* x, const float* e, int64_t total, hipStream_t s) {
    hipLaunchKernelGGL(axpy_correction_kernel, dim3(gridFor(total)), dim3(256), 0, s, x, e, total);
}
void launch_fill_one(void* p, int64_t n, bool fp64, hipStream_t s) {
    if (fp64) hipLaunchKernelGGL(fill_one_kernel<double>, dim3(gridFor(n)), dim3(256), 0, s, (double*)p, n);
    else hipLaunchKernelGGL(fill_one_kernel<float>, dim3(gridFor(n)), dim3(256), 0, s, (float*)p, n);
}
void launch_zero(void* p, int64_t bytes, hipStream_t s) {
    hipLaunchKernelGGL(zero_kernel, dim3(gridFor(bytes / 4)), dim3(256), 0, s, p, bytes);
}
void launch_zero_rows(double* a, int64_t n, const int* rows, int nrows, hipStream_t s) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(zero_rows_kernel, dim3(std::min<unsigned>(gridFor(n), 64u), nrows), dim3(256), 0, s, a, n, rows);
}
int cg_partial_blocks(int64_t n) { return (int)gridFor(n); }

#define LSG_INST(V)                                                                                                  \
    template void launch_nodal<V>(const NodalItem*, int, int, const V*, int64_t, V*, int64_t, int64_t, int,         \
                                  const CGState*, hipStream_t, const int*);                                              \
    template void launch_prol<V>(const ProlItem*, int, int, const V*, int64_t, V*, int64_t, int64_t, int64_t, int,      \
                                 const CGState*,                                                                     \
                                 hipStream_t, const int*);                                                           \
    template int launch_apply<V>(const ApplyArgs<V>&, int, int, int, size_t, hipStream_t);                          \
    template void launch_sum<V>(const SumItem*, int, const int64_t*, int, const V*, int64_t, V*, int64_t, int,       \
                                const CGState*, hipStream_t, const int*);                                                \
    template void launch_diag<V>(const SumItem*, int, int, int, int, const V*, int64_t, int64_t, V*, int64_t, int,   \
                                 hipStream_t);                                                                       \
    template void launch_rhs_sines<V>(const SumItem*, int, int, int, const double*, const int32_t*, V*, int64_t,     \
                                      int, hipStream_t, const int32_t*, const uint8_t*);                             \
    template void launch_rhs_tab<V>(const SumItem*, int, int, int, const double*, const double*, int64_t, V*,        \
                                    int64_t, int, hipStream_t, const int32_t*, const uint8_t*);                      \
    template void launch_cg_init<V>(const V*, V*, V*, V*, const V*, int64_t, int, double*, int, hipStream_t,          \
                                    const BlockPreArgs<V>&);                                                        \
    template void launch_cg_iter_tail<V>(V*, void*, bool, V*, V*, const V*, const V*, int64_t, int, double*, int,     \
                                         CGState*, hipStream_t, const BlockPreArgs<V>&);
LSG_INST(double)
LSG_INST(float)
#undef LSG_INST

}  // namespace lsg
