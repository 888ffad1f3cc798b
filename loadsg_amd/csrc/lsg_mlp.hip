// lsg_mlp.hip -- warm-start network for the batched solves: an MLP
// n_in -> width (x hidden layers, GELU) -> n_out mapping the parameters of a
// problem instance (coefficient and right-hand-side amplitudes) to an initial guess
// of its solution.  No counterpart in LoAdSG.
//
// GEMMs run on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16): BF16 operands,
// FP32 accumulation.  Master weights, gradients and Adam moments are FP32; the
// forward pass keeps BF16 copies of the weights and activations as MFMA operands.
// Every reduction has a fixed order (split-K partials summed in split order,
// column sums row by row, loss partials block by block), so results are
// bit-reproducible for a given batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/loadsg_amd.h"

namespace lsg {
int api_fail(int code, const std::string& msg);
}

namespace {

using bf16 = __bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

#define MLPCHK(x)                                                                                              \
    do {                                                                                                       \
        hipError_t e_ = (x);                                                                                   \
        if (e_ != hipSuccess)                                                                                  \
            return lsg::api_fail(LSG_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_));             \
    } while (0)

__device__ inline float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// ------------------------------------------------------------------ GEMM
// C[m][n] = sum_k A(m,k) B(k,n), A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].
// Four waves in a 2 x 2 arrangement; each wave computes WM x WN tiles of 32 x 32 with
// 32x32x16 BF16 MFMAs (FP32 accumulation), so the workgroup tile is 64 WM x 64 WN.
// K steps of 64.  Tiles are staged in LDS with K contiguous (B transposed) so that every
// lane reads its 8 BF16 operands with one 16-byte LDS load; the global loads put
// consecutive lanes on consecutive addresses of the contiguous dimension.  blockIdx.z
// splits K: split z covers [z*kChunk, (z+1)*kChunk) and writes its partial sum to
// C + z*cSplit.
constexpr int BK = 64, LDK = BK + 8;

struct GemmArgs {
    const bf16* A;
    int64_t sam, sak;
    const bf16* B;
    int64_t sbk, sbn;
    int M, N, K;
    int kChunk;
    const float* bias;   // [N] or null
    float* C;            // FP32 result (pre-activation) or null
    int64_t ldc, cSplit;
    bf16* Cb;            // BF16 result (GELU applied if act) or null
    int64_t ldcb;
    int act;
};

// One operand tile (T rows of the tile dimension x BK of K) staged into LDS as
// dst[row * LDK + k].  Each thread owns fixed rows / k offsets for the whole K loop;
// its loads walk a pointer by a fixed increment, so the per-load work is one add and
// one bounds test.
template <int T>
struct TileLoader {
    static constexpr int NL = T * BK / 256;   // loads per thread per K step
    const bf16* base;        // address of this thread's first element at k = kBeg
    int64_t step;            // address increment per K step of BK
    int64_t inc;             // address increment between this thread's loads
    int dst0, dinc;          // LDS index of the first load and its increment
    int kof0, kinc;          // k offset (inside the step) of the first load and its increment
    int r0, rinc, t0, nt;    // tile row of the first load, its increment, tile origin, rows
    __device__ TileLoader(const bf16* P, int64_t st, int64_t sk, int t0_, int nt_, int kBeg) : t0(t0_), nt(nt_) {
        const int t = threadIdx.x;
        if (sk == 1) {       // K contiguous: lane -> k = t % BK, rows t / BK + (256 / BK) i
            const int kk = t % BK;
            r0 = t / BK;
            rinc = 256 / BK;
            base = P + (int64_t)(t0 + r0) * st + kBeg + kk;
            step = BK;
            inc = (int64_t)rinc * st;
            dst0 = r0 * LDK + kk;
            dinc = rinc * LDK;
            kof0 = kk;
            kinc = 0;
        } else {             // tile dimension contiguous: lane -> row t % T, k = t / T + (256 / T) i
            const int k0 = t / T;
            r0 = t % T;
            rinc = 0;
            base = P + (int64_t)(t0 + r0) * st + (int64_t)(kBeg + k0) * sk;
            step = (int64_t)BK * sk;
            inc = (int64_t)(256 / T) * sk;
            dst0 = r0 * LDK + k0;
            dinc = 256 / T;
            kof0 = k0;
            kinc = 256 / T;
        }
    }
    // global -> registers for the K step starting at p (kRem values of K left)
    __device__ void fetch(bf16 (&v)[NL], const bf16* p, int kRem) const {
        const bf16 zero = bf16(0.0f);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
            const bool in = t0 + r0 + rinc * i < nt && kof0 + kinc * i < kRem;
            v[i] = in ? p[inc * i] : zero;
        }
    }
    __device__ void store(bf16* __restrict__ lds, const bf16 (&v)[NL]) const {
#pragma unroll
        for (int i = 0; i < NL; ++i) lds[dst0 + dinc * i] = v[i];
    }
};

template <int WM, int WN>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
    constexpr int TM = 64 * WM, TN = 64 * WN;
    // two LDS buffers per operand: step k is multiplied from one while step k + 1 is
    // stored into the other, one barrier per K step
    __shared__ __attribute__((aligned(16))) bf16 As2[2][TM * LDK];
    __shared__ __attribute__((aligned(16))) bf16 Bs2[2][TN * LDK];
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kBeg = blockIdx.z * g.kChunk;
    const int kEnd = min(g.K, kBeg + g.kChunk);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    floatx16 acc[WM][WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;
    const TileLoader<TM> la(g.A, g.sam, g.sak, m0, g.M, kBeg);
    const TileLoader<TN> lb(g.B, g.sbn, g.sbk, n0, g.N, kBeg);
    const bf16* pa = la.base;
    const bf16* pb = lb.base;
    // the next K step is fetched into registers while this one is multiplied
    bf16 ra[TileLoader<TM>::NL], rb[TileLoader<TN>::NL];
    if (kBeg < kEnd) {
        la.fetch(ra, pa, kEnd - kBeg);
        lb.fetch(rb, pb, kEnd - kBeg);
        la.store(As2[0], ra);
        lb.store(Bs2[0], rb);
    }
    __syncthreads();
    int cur = 0;
    for (int k0 = kBeg; k0 < kEnd; k0 += BK) {
        const bool more = k0 + BK < kEnd;
        if (more) {
            pa += la.step;
            pb += lb.step;
            la.fetch(ra, pa, kEnd - k0 - BK);
            lb.fetch(rb, pb, kEnd - k0 - BK);
        }
        const bf16* As = As2[cur];
        const bf16* Bs = Bs2[cur];
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            bf16x8 a[WM], b[WN];
#pragma unroll
            for (int i = 0; i < WM; ++i)
                a[i] = *reinterpret_cast<const bf16x8*>(
                    &As[(wm * 32 * WM + i * 32 + (lane & 31)) * LDK + ks + 8 * (lane >> 5)]);
#pragma unroll
            for (int j = 0; j < WN; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(
                    &Bs[(wn * 32 * WN + j * 32 + (lane & 31)) * LDK + ks + 8 * (lane >> 5)]);
#pragma unroll
            for (int i = 0; i < WM; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            la.store(As2[cur ^ 1], ra);
            lb.store(Bs2[cur ^ 1], rb);
        }
        __syncthreads();
        cur ^= 1;
    }
    // 32x32 result layout: lane holds column (lane & 31), rows 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + wm * 32 * WM + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int n = n0 + wn * 32 * WN + j * 32 + (lane & 31);
                if (m < g.M && n < g.N) {
                    float v = acc[i][j][r];
                    if (g.bias) v += g.bias[n];
                    if (g.C) g.C[(int64_t)blockIdx.z * g.cSplit + (int64_t)m * g.ldc + n] = v;
                    if (g.Cb) g.Cb[(int64_t)m * g.ldcb + n] = bf16(g.act ? gelu_f(v) : v);
                }
            }
}

// sum of the split-K partials in split order (+ bias, GELU, BF16 copy); C is M x N dense
__global__ void __launch_bounds__(256) split_reduce_kernel(const float* __restrict__ part, int splits, int64_t stride,
                                                           int M, int N, const float* __restrict__ bias,
                                                           float* __restrict__ C, bf16* __restrict__ Cb, int act) {
    const int64_t total = (int64_t)M * N;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        float v = 0.0f;
        for (int s = 0; s < splits; ++s) v += part[s * stride + i];
        if (bias) v += bias[i % N];
        if (C) C[i] = v;
        if (Cb) Cb[i] = bf16(act ? gelu_f(v) : v);
    }
}

// ------------------------------------------------------------ elementwise
__global__ void to_bf16_kernel(const float* __restrict__ x, bf16* __restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = bf16(x[i]);
}

// dY = 2 (Y - T) / count (mean squared error), FP32 and BF16; per-block partial
// sums of (Y - T)^2 for the loss
__global__ void __launch_bounds__(256) loss_grad_kernel(const float* __restrict__ y, const float* __restrict__ tgt,
                                                        int64_t n, float scale, float* __restrict__ dyf,
                                                        bf16* __restrict__ dyb, double* __restrict__ partials) {
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float e = y[i] - tgt[i];
        s += (double)e * (double)e;
        const float g = scale * e;
        dyf[i] = g;
        dyb[i] = bf16(g);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// dZ = dA * gelu'(Z), FP32 and BF16
__global__ void gelu_back_kernel(const float* __restrict__ da, const float* __restrict__ z, float* __restrict__ dzf,
                                 bf16* __restrict__ dzb, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float g = da[i] * gelu_grad(z[i]);
        dzf[i] = g;
        dzb[i] = bf16(g);
    }
}

// out[n] = sum over rows m of X[m][n]: 64 columns per workgroup, 4 row slices of
// consecutive rows per column summed in order, then the slices in order (fixed order,
// deterministic)
__global__ void __launch_bounds__(256) colsum_kernel(const float* __restrict__ X, int M, int N,
                                                     float* __restrict__ out) {
    __shared__ float part[4][64];
    const int c = threadIdx.x & 63, sl = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    const int rows = (M + 3) / 4, r0 = sl * rows, r1 = min(M, r0 + rows);
    float s = 0.0f;
    if (n < N)
        for (int m = r0; m < r1; ++m) s += X[(int64_t)m * N + n];
    part[sl][c] = s;
    __syncthreads();
    if (sl == 0 && n < N) out[n] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

// out[b][i] = sum_j w[b*J + j] U[(rows[b]*J + j)*n + i], j in order
__global__ void __launch_bounds__(256) mix_rows_kernel(const float* __restrict__ U, int64_t n, int J,
                                                       const int64_t* __restrict__ rows, const float* __restrict__ w,
                                                       float* __restrict__ out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t b = blockIdx.y;
    const float* u = U + rows[b] * J * n + i;
    const float* wb = w + b * J;
    float s = 0.0f;
    for (int j = 0; j < J; ++j) s = fmaf(wb[j], u[(int64_t)j * n], s);
    out[b * n + i] = s;
}

// Adam with bias correction; refreshes the BF16 copy of the weights
__global__ void adam_kernel(float* __restrict__ w, bf16* __restrict__ wb, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = bf16(wi);
    }
}

int gridFor(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 4096); }

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t b) {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = std::max<size_t>(b, 16);
        return hipMalloc(&p, bytes);
    }
    template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
};

struct Layer {
    int in = 0, out = 0;
    Buf W, Wb, b, gW, gb, mW, vW, mb, vb;   // W is in x out, row-major
};

}  // namespace

struct lsg_mlp {
    int nIn = 0, width = 0, hidden = 0;
    int64_t nOut = 0;
    std::vector<std::unique_ptr<Layer>> layers;
    int cap = 0;                      // batch capacity of the activation buffers
    Buf Xb;                           // BF16 input
    std::vector<std::unique_ptr<Buf>> Z, Ab;   // per hidden layer: FP32 pre-activation, BF16 activation
    Buf Y, dYf, dYb, dA, dZf, dZb, part, lossPart;
    int64_t step = 0;
    hipStream_t s = nullptr;
    ~lsg_mlp() { if (s) (void)hipStreamDestroy(s); }
};

namespace {

constexpr int SPLIT_K_TARGET = 4096;   // K per split for long reductions

int gemm(lsg_mlp* m, const bf16* A, int64_t sam, int64_t sak, const bf16* B, int64_t sbk, int64_t sbn, int M, int N,
         int K, const float* bias, float* C, bf16* Cb, int act) {
    // 64 x 64 workgroup tiles: measured faster here than 128 x 128 (2 x 2 MFMA tiles per
    // wave) and than K steps of 128
    constexpr int TM = 64, TN = 64;
    const int64_t tiles = (int64_t)((M + TM - 1) / TM) * ((N + TN - 1) / TN);
    int splits = 1;
    if (K > 2 * SPLIT_K_TARGET) splits = (K + SPLIT_K_TARGET - 1) / SPLIT_K_TARGET;
    if (tiles * splits < 256 && K >= 4 * BK) {
        // few workgroups: split K further so that about 256 run, at least 2 K steps each
        splits = (int)std::max<int64_t>(splits, std::min<int64_t>((256 + tiles - 1) / tiles, K / (2 * BK)));
    }
    GemmArgs g{};
    g.A = A; g.sam = sam; g.sak = sak;
    g.B = B; g.sbk = sbk; g.sbn = sbn;
    g.M = M; g.N = N; g.K = K;
    g.act = act;
    const dim3 blk(256);
    auto launch = [&](int zs) {
        const dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, zs);
        hipLaunchKernelGGL((gemm_bf16_kernel<1, 1>), grid, blk, 0, m->s, g);
    };
    if (splits == 1) {
        g.kChunk = ((K + BK - 1) / BK) * BK;
        g.bias = bias; g.C = C; g.ldc = N; g.cSplit = 0; g.Cb = Cb; g.ldcb = N;
        launch(1);
    } else {
        g.kChunk = ((K / splits + BK - 1) / BK + 1) * BK;
        const int used = (K + g.kChunk - 1) / g.kChunk;
        const int64_t need = (int64_t)used * M * N * 4;
        if ((int64_t)m->part.bytes < need) MLPCHK(m->part.alloc(need));
        g.bias = nullptr; g.C = m->part.as<float>(); g.ldc = N; g.cSplit = (int64_t)M * N; g.Cb = nullptr;
        launch(used);
        hipLaunchKernelGGL(split_reduce_kernel, dim3(gridFor((int64_t)M * N)), blk, 0, m->s, m->part.as<float>(), used,
                           (int64_t)M * N, M, N, bias, C, Cb, act);
    }
    MLPCHK(hipGetLastError());
    return LSG_OK;
}

int ensureCap(lsg_mlp* m, int batch) {
    if (batch <= m->cap) return LSG_OK;
    const int64_t B = batch;
    MLPCHK(m->Xb.alloc(B * m->nIn * 2));
    m->Z.clear();
    m->Ab.clear();
    for (int l = 0; l < m->hidden; ++l) {
        m->Z.emplace_back(new Buf());
        m->Ab.emplace_back(new Buf());
        MLPCHK(m->Z.back()->alloc(B * m->width * 4));
        MLPCHK(m->Ab.back()->alloc(B * m->width * 2));
    }
    MLPCHK(m->Y.alloc(B * m->nOut * 4));
    MLPCHK(m->dYf.alloc(B * m->nOut * 4));
    MLPCHK(m->dYb.alloc(B * m->nOut * 2));
    MLPCHK(m->dA.alloc(B * m->width * 4));
    MLPCHK(m->dZf.alloc(B * m->width * 4));
    MLPCHK(m->dZb.alloc(B * m->width * 2));
    MLPCHK(m->lossPart.alloc(4096 * 8));
    m->cap = batch;
    return LSG_OK;
}

// forward pass into m->Y (FP32, batch x nOut) or y
int forward(lsg_mlp* m, const float* x, int batch, float* y) {
    int rc = ensureCap(m, batch);
    if (rc) return rc;
    hipLaunchKernelGGL(to_bf16_kernel, dim3(gridFor((int64_t)batch * m->nIn)), dim3(256), 0, m->s, x,
                       m->Xb.as<bf16>(), (int64_t)batch * m->nIn);
    const bf16* a = m->Xb.as<bf16>();
    for (int l = 0; l <= m->hidden; ++l) {
        Layer& L = *m->layers[l];
        const bool last = l == m->hidden;
        rc = gemm(m, a, L.in, 1, L.Wb.as<bf16>(), L.out, 1, batch, L.out, L.in, L.b.as<float>(),
                  last ? y : m->Z[l]->as<float>(), last ? nullptr : m->Ab[l]->as<bf16>(), last ? 0 : 1);
        if (rc) return rc;
        if (!last) a = m->Ab[l]->as<bf16>();
    }
    return LSG_OK;
}

// loss and gradients of the mean squared error against tgt
int gradients(lsg_mlp* m, const float* x, const float* tgt, int batch, double* loss) {
    int rc = ensureCap(m, batch);
    if (rc) return rc;
    rc = forward(m, x, batch, m->Y.as<float>());
    if (rc) return rc;
    const int64_t ny = (int64_t)batch * m->nOut;
    const int nb = gridFor(ny);
    hipLaunchKernelGGL(loss_grad_kernel, dim3(nb), dim3(256), 0, m->s, m->Y.as<float>(), tgt, ny, 2.0f / float(ny),
                       m->dYf.as<float>(), m->dYb.as<bf16>(), m->lossPart.as<double>());
    const float* dZf = m->dYf.as<float>();
    const bf16* dZb = m->dYb.as<bf16>();
    for (int l = m->hidden; l >= 0; --l) {
        Layer& L = *m->layers[l];
        const bf16* aPrev = l > 0 ? m->Ab[l - 1]->as<bf16>() : m->Xb.as<bf16>();
        // dW = A_prev^T dZ  (in x out, K = batch)
        rc = gemm(m, aPrev, 1, L.in, dZb, L.out, 1, L.in, L.out, batch, nullptr, L.gW.as<float>(), nullptr, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(colsum_kernel, dim3((L.out + 63) / 64), dim3(256), 0, m->s, dZf, batch, L.out,
                           L.gb.as<float>());
        if (l == 0) break;
        // dA_prev = dZ W^T  (batch x in, K = out)
        rc = gemm(m, dZb, L.out, 1, L.Wb.as<bf16>(), 1, L.out, batch, L.in, L.out, nullptr, m->dA.as<float>(),
                  nullptr, 0);
        if (rc) return rc;
        const int64_t nz = (int64_t)batch * L.in;
        hipLaunchKernelGGL(gelu_back_kernel, dim3(gridFor(nz)), dim3(256), 0, m->s, m->dA.as<float>(),
                           m->Z[l - 1]->as<float>(), m->dZf.as<float>(), m->dZb.as<bf16>(), nz);
        dZf = m->dZf.as<float>();
        dZb = m->dZb.as<bf16>();
    }
    MLPCHK(hipGetLastError());
    if (loss) {
        std::vector<double> hp(nb);
        MLPCHK(hipMemcpyAsync(hp.data(), m->lossPart.p, nb * 8, hipMemcpyDeviceToHost, m->s));
        MLPCHK(hipStreamSynchronize(m->s));
        double sum = 0.0;
        for (double v : hp) sum += v;
        *loss = sum / double(ny);
    }
    return LSG_OK;
}

// deterministic normal samples (splitmix64 + Box-Muller)
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};

}  // namespace

extern "C" {

int lsg_mlp_create(int n_in, int width, int hidden, int64_t n_out, uint64_t seed, lsg_mlp** out) {
    if (!out || n_in < 1 || width < 1 || hidden < 1 || n_out < 1 || n_out > (int64_t(1) << 30))
        return lsg::api_fail(LSG_ERR_ARG, "bad MLP shape");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return lsg::api_fail(LSG_ERR_DEVICE, "no GPU");
    std::unique_ptr<lsg_mlp> m(new lsg_mlp());
    m->nIn = n_in;
    m->width = width;
    m->hidden = hidden;
    m->nOut = n_out;
    MLPCHK(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking));
    Rng rng{seed};
    for (int l = 0; l <= hidden; ++l) {
        std::unique_ptr<Layer> L(new Layer());
        L->in = l == 0 ? n_in : width;
        L->out = l == hidden ? (int)n_out : width;
        const int64_t nw = (int64_t)L->in * L->out;
        // He initialisation for the GELU layers, 1/sqrt(in) for the linear output layer
        const double sd = (l == hidden ? 1.0 : std::sqrt(2.0)) / std::sqrt((double)L->in);
        std::vector<float> w(nw);
        for (auto& v : w) v = float(sd * rng.normal());
        MLPCHK(L->W.alloc(nw * 4));
        MLPCHK(L->Wb.alloc(nw * 2));
        MLPCHK(L->gW.alloc(nw * 4));
        MLPCHK(L->mW.alloc(nw * 4));
        MLPCHK(L->vW.alloc(nw * 4));
        MLPCHK(L->b.alloc((size_t)L->out * 4));
        MLPCHK(L->gb.alloc((size_t)L->out * 4));
        MLPCHK(L->mb.alloc((size_t)L->out * 4));
        MLPCHK(L->vb.alloc((size_t)L->out * 4));
        MLPCHK(hipMemcpy(L->W.p, w.data(), nw * 4, hipMemcpyHostToDevice));
        MLPCHK(hipMemset(L->mW.p, 0, nw * 4));
        MLPCHK(hipMemset(L->vW.p, 0, nw * 4));
        MLPCHK(hipMemset(L->b.p, 0, (size_t)L->out * 4));
        MLPCHK(hipMemset(L->mb.p, 0, (size_t)L->out * 4));
        MLPCHK(hipMemset(L->vb.p, 0, (size_t)L->out * 4));
        hipLaunchKernelGGL(to_bf16_kernel, dim3(gridFor(nw)), dim3(256), 0, m->s, L->W.as<float>(), L->Wb.as<bf16>(),
                           nw);
        m->layers.push_back(std::move(L));
    }
    MLPCHK(hipStreamSynchronize(m->s));
    *out = m.release();
    return LSG_OK;
}

int lsg_mlp_forward(lsg_mlp* m, const float* x, float* y, int batch) {
    if (!m || !x || !y || batch < 1) return lsg::api_fail(LSG_ERR_ARG, "bad arguments");
    int rc = forward(m, x, batch, y);
    if (rc) return rc;
    MLPCHK(hipStreamSynchronize(m->s));
    return LSG_OK;
}

int lsg_mlp_gradients(lsg_mlp* m, const float* x, const float* target, int batch, double* loss) {
    if (!m || !x || !target || batch < 1) return lsg::api_fail(LSG_ERR_ARG, "bad arguments");
    int rc = gradients(m, x, target, batch, loss);
    if (rc) return rc;
    MLPCHK(hipStreamSynchronize(m->s));
    return LSG_OK;
}

int lsg_mlp_train_step(lsg_mlp* m, const float* x, const float* target, int batch, float lr, double* loss) {
    if (!m || !x || !target || batch < 1 || !(lr > 0.0f)) return lsg::api_fail(LSG_ERR_ARG, "bad arguments");
    int rc = gradients(m, x, target, batch, loss);
    if (rc) return rc;
    m->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
    for (auto& Lp : m->layers) {
        Layer& L = *Lp;
        const int64_t nw = (int64_t)L.in * L.out;
        hipLaunchKernelGGL(adam_kernel, dim3(gridFor(nw)), dim3(256), 0, m->s, L.W.as<float>(), L.Wb.as<bf16>(),
                           L.gW.as<float>(), L.mW.as<float>(), L.vW.as<float>(), nw, lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(adam_kernel, dim3(gridFor(L.out)), dim3(256), 0, m->s, L.b.as<float>(), (bf16*)nullptr,
                           L.gb.as<float>(), L.mb.as<float>(), L.vb.as<float>(), (int64_t)L.out, lr, b1, b2, eps, c1,
                           c2);
    }
    MLPCHK(hipGetLastError());
    MLPCHK(hipStreamSynchronize(m->s));
    return LSG_OK;
}

int lsg_mlp_num_params(const lsg_mlp* m, int64_t* n) {
    if (!m || !n) return lsg::api_fail(LSG_ERR_ARG, "null argument");
    int64_t t = 0;
    for (auto& L : m->layers) t += (int64_t)L->in * L->out + L->out;
    *n = t;
    return LSG_OK;
}

// flat order: for each layer W (in x out, row-major) then b (out)
static int mlpCopy(const lsg_mlp* m, float* host, bool grads, bool toHost) {
    int64_t o = 0;
    for (auto& Lp : m->layers) {
        Layer& L = *Lp;
        const int64_t nw = (int64_t)L.in * L.out;
        void* w = grads ? L.gW.p : L.W.p;
        void* b = grads ? L.gb.p : L.b.p;
        if (toHost) {
            MLPCHK(hipMemcpy(host + o, w, nw * 4, hipMemcpyDeviceToHost));
            MLPCHK(hipMemcpy(host + o + nw, b, (size_t)L.out * 4, hipMemcpyDeviceToHost));
        } else {
            MLPCHK(hipMemcpy(w, host + o, nw * 4, hipMemcpyHostToDevice));
            MLPCHK(hipMemcpy(b, host + o + nw, (size_t)L.out * 4, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(to_bf16_kernel, dim3(gridFor(nw)), dim3(256), 0, m->s, L.W.as<float>(),
                               L.Wb.as<bf16>(), nw);
        }
        o += nw + L.out;
    }
    MLPCHK(hipStreamSynchronize(m->s));
    return LSG_OK;
}

int lsg_mlp_get_params(const lsg_mlp* m, float* host) {
    if (!m || !host) return lsg::api_fail(LSG_ERR_ARG, "null argument");
    return mlpCopy(m, host, false, true);
}

int lsg_mlp_set_params(lsg_mlp* m, const float* host) {
    if (!m || !host) return lsg::api_fail(LSG_ERR_ARG, "null argument");
    return mlpCopy(m, const_cast<float*>(host), false, false);
}

int lsg_mlp_get_grads(const lsg_mlp* m, float* host) {
    if (!m || !host) return lsg::api_fail(LSG_ERR_ARG, "null argument");
    return mlpCopy(m, host, true, true);
}

int lsg_mlp_mix_rows(lsg_mlp* m, const float* U, int64_t n, int J, const int64_t* rows, const float* w, int batch,
                     float* out) {
    if (!m || !U || !rows || !w || !out || n < 1 || J < 1 || batch < 1 || batch > 65535)
        return lsg::api_fail(LSG_ERR_ARG, "bad mix_rows arguments");
    hipLaunchKernelGGL(mix_rows_kernel, dim3((unsigned)((n + 255) / 256), batch), dim3(256), 0, m->s, U, n, J, rows, w,
                       out);
    MLPCHK(hipGetLastError());
    MLPCHK(hipStreamSynchronize(m->s));
    return LSG_OK;
}

int lsg_mlp_destroy(lsg_mlp* m) {
    delete m;
    return LSG_OK;
}

}  // extern "C"
