// MLP surrogate kernels: BF16 GEMMs on the CDNA4 matrix cores (v_mfma_f32_32x32x16_bf16, FP32
// accumulation) with fused epilogues, the MSE loss, bias gradients, minibatch gather and Adam.
//
// GEMM: C[i][j] = sum_k A(i, k) B(k, j), A(i, k) = A[i * lda + k] (A_T = false) or A[k * lda + i]
// (A_T = true), B(k, j) = B[k * ldb + j] (B_T = false) or B[j * ldb + k] (B_T = true).
// Workgroup tile 64 x 64 (one MFMA tile per wavefront) or 128 x 128 (2 x 2 per wavefront), four
// wavefronts, K in steps of 32 staged through LDS with the k index contiguous per row (so every
// MFMA operand is one 16-byte LDS read per lane). Global
// loads are 16 B along whichever index is contiguous in memory. Rows and columns past M / N and
// k past K read as zero.
//
// Operand layout of the 32x32x16 MFMA (64 lanes): lane l supplies A[l % 32][8 (l / 32) + e] and
// B[8 (l / 32) + e][l % 32] for e = 0..7, and receives D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32]
// in accumulator r = 0..15. A and B use the same k order, so the sum does not depend on it.
#include "internal.h"

namespace emfd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int BK = 32, LDK = BK + 8;  // LDS row stride (bf16): 80 B, 16 B aligned

__device__ __forceinline__ float gelu_f(float x) {
    // tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))
    const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad_f(float x) {
    const float x2 = x * x;
    const float u = 0.7978845608028654f * (x + 0.044715f * x2 * x);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x2);
}

struct GemmArgs {
    const __bf16 *A;
    const __bf16 *B;
    int M, N, K;
    int lda, ldb;
    // epilogue
    int epi;              // MLP_EPI_*
    const float *bias;    // [N]
    const float *aux;     // pre-activation Z [M][N] (gelu-grad epilogue)
    float *out_f32;       // [M][ldc]
    __bf16 *out_bf16;     // [M][ldc]
    int ldc;
};

enum : int { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_GRAD = 3, EPI_BIAS_GELU_INFER = 4 };

template <bool A_T, bool B_T, int W>
__device__ __forceinline__ void stage(const GemmArgs &g, int i0, int j0, int k0, __bf16 (*sA)[LDK],
                                      __bf16 (*sB)[LDK]) {
    const int t = threadIdx.x;
    const __bf16 zero = static_cast<__bf16>(0.0f);
#pragma unroll
    for (int e = 0; e < W; ++e) {
        // A tile: rows i0 .. i0 + 64 W - 1, k0 .. k0 + 31 -> sA[i][k]
        if (!A_T) {
            const int i = (t >> 2) + 64 * e, kk = (t & 3) * 8;
            const int gi = i0 + i, gk = k0 + kk;
            __bf16 v[8];
            if (gi < g.M && gk + 8 <= g.K && (g.lda % 8) == 0) {
                *reinterpret_cast<uint4 *>(v) =
                    *reinterpret_cast<const uint4 *>(g.A + static_cast<size_t>(gi) * g.lda + gk);
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    v[q] = (gi < g.M && gk + q < g.K) ? g.A[static_cast<size_t>(gi) * g.lda + gk + q] : zero;
            }
            *reinterpret_cast<uint4 *>(&sA[i][kk]) = *reinterpret_cast<const uint4 *>(v);
        } else {
            const int kk = t >> 3, i = (t & 7) * 8 + 64 * e;
            const int gk = k0 + kk, gi = i0 + i;
            __bf16 v[8];
            if (gk < g.K && gi + 8 <= g.M && (g.lda % 8) == 0) {
                *reinterpret_cast<uint4 *>(v) =
                    *reinterpret_cast<const uint4 *>(g.A + static_cast<size_t>(gk) * g.lda + gi);
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    v[q] = (gk < g.K && gi + q < g.M) ? g.A[static_cast<size_t>(gk) * g.lda + gi + q] : zero;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) sA[i + q][kk] = v[q];
        }
        // B tile: k0 .. k0 + 31, cols j0 .. j0 + 64 W - 1 -> sB[j][k]
        if (B_T) {
            const int j = (t >> 2) + 64 * e, kk = (t & 3) * 8;
            const int gj = j0 + j, gk = k0 + kk;
            __bf16 v[8];
            if (gj < g.N && gk + 8 <= g.K && (g.ldb % 8) == 0) {
                *reinterpret_cast<uint4 *>(v) =
                    *reinterpret_cast<const uint4 *>(g.B + static_cast<size_t>(gj) * g.ldb + gk);
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    v[q] = (gj < g.N && gk + q < g.K) ? g.B[static_cast<size_t>(gj) * g.ldb + gk + q] : zero;
            }
            *reinterpret_cast<uint4 *>(&sB[j][kk]) = *reinterpret_cast<const uint4 *>(v);
        } else {
            const int kk = t >> 3, j = (t & 7) * 8 + 64 * e;
            const int gk = k0 + kk, gj = j0 + j;
            __bf16 v[8];
            if (gk < g.K && gj + 8 <= g.N && (g.ldb % 8) == 0) {
                *reinterpret_cast<uint4 *>(v) =
                    *reinterpret_cast<const uint4 *>(g.B + static_cast<size_t>(gk) * g.ldb + gj);
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    v[q] = (gk < g.K && gj + q < g.N) ? g.B[static_cast<size_t>(gk) * g.ldb + gj + q] : zero;
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) sB[j + q][kk] = v[q];
        }
    }
}

// Workgroup tile (64 W) x (64 W); each wavefront owns (32 W) x (32 W) = W x W MFMA tiles.
template <bool A_T, bool B_T, int W>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
    __shared__ __attribute__((aligned(16))) __bf16 sA[64 * W][LDK];
    __shared__ __attribute__((aligned(16))) __bf16 sB[64 * W][LDK];
    const int i0 = blockIdx.y * 64 * W, j0 = blockIdx.x * 64 * W;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32 * W, wc = (wave & 1) * 32 * W;
    floatx16 acc[W][W];
#pragma unroll
    for (int tm = 0; tm < W; ++tm)
#pragma unroll
        for (int tn = 0; tn < W; ++tn) acc[tm][tn] = floatx16{};
    for (int k0 = 0; k0 < g.K; k0 += BK) {
        stage<A_T, B_T, W>(g, i0, j0, k0, sA, sB);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK; ks += 16) {
            bf16x8 a[W], b[W];
#pragma unroll
            for (int tm = 0; tm < W; ++tm)
                a[tm] = *reinterpret_cast<const bf16x8 *>(&sA[wr + 32 * tm + (lane & 31)][ks + 8 * (lane >> 5)]);
#pragma unroll
            for (int tn = 0; tn < W; ++tn)
                b[tn] = *reinterpret_cast<const bf16x8 *>(&sB[wc + 32 * tn + (lane & 31)][ks + 8 * (lane >> 5)]);
#pragma unroll
            for (int tm = 0; tm < W; ++tm)
#pragma unroll
                for (int tn = 0; tn < W; ++tn)
                    acc[tm][tn] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[tm], b[tn], acc[tm][tn], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int tn = 0; tn < W; ++tn) {
        const int j = j0 + wc + 32 * tn + (lane & 31);
        if (j >= g.N) continue;
        const float bj = (g.epi == EPI_BIAS || g.epi == EPI_BIAS_GELU || g.epi == EPI_BIAS_GELU_INFER) ? g.bias[j] : 0.0f;
#pragma unroll
        for (int tm = 0; tm < W; ++tm)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = i0 + wr + 32 * tm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (i >= g.M) continue;
                const size_t o = static_cast<size_t>(i) * g.ldc + j;
                float v = acc[tm][tn][r];
                if (g.epi == EPI_BIAS || g.epi == EPI_BIAS_GELU || g.epi == EPI_BIAS_GELU_INFER) v += bj;
                if (g.epi == EPI_BIAS_GELU_INFER) {
                    g.out_bf16[o] = static_cast<__bf16>(gelu_f(v));
                } else if (g.epi == EPI_BIAS_GELU) {
                    g.out_f32[o] = v;  // pre-activation, kept for the backward pass
                    g.out_bf16[o] = static_cast<__bf16>(gelu_f(v));
                } else if (g.epi == EPI_GELU_GRAD) {
                    const float d = v * gelu_grad_f(g.aux[o]);
                    g.out_f32[o] = d;
                    g.out_bf16[o] = static_cast<__bf16>(d);
                } else {
                    if (g.out_f32) g.out_f32[o] = v;
                    if (g.out_bf16) g.out_bf16[o] = static_cast<__bf16>(v);
                }
            }
    }
}

// MSE loss over M x N outputs: grad = 2 (y - t) / (M N) (FP32 and BF16), per-block partial sums
// of (y - t)^2 in FP64.
__global__ void __launch_bounds__(256) k_mse(const float *y, const float *t, size_t n, float scale, float *g,
                                             __bf16 *gb, double *partials) {
    __shared__ double s[256];
    double acc = 0.0;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256) {
        const float d = y[i] - t[i];
        acc += static_cast<double>(d) * d;
        const float gi = scale * d;
        g[i] = gi;
        gb[i] = static_cast<__bf16>(gi);
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = s[0];
}

// db[j] = sum_i g[i][j]
__global__ void __launch_bounds__(256) k_colsum(const float *g, int M, int N, float *db) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= N) return;
    float acc = 0.0f;
    for (int i = 0; i < M; ++i) acc += g[static_cast<size_t>(i) * N + j];
    db[j] = acc;
}

// rows idx[r] of x (FP32, width k_in) -> BF16 rows of width ld (zero padded), and of t (FP32, width n_t)
__global__ void __launch_bounds__(256) k_gather(const float *x, int k_in, int ld, const float *t, int n_t,
                                                const int *idx, int rows, __bf16 *xb, float *tb) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const size_t src = static_cast<size_t>(idx ? idx[r] : r);
    for (int c = threadIdx.x; c < ld; c += 256)
        xb[static_cast<size_t>(r) * ld + c] = static_cast<__bf16>(c < k_in ? x[src * k_in + c] : 0.0f);
    if (t)
        for (int c = threadIdx.x; c < n_t; c += 256) tb[static_cast<size_t>(r) * n_t + c] = t[src * n_t + c];
}

// Adam on FP32 master parameters; refreshes the BF16 weight copies (bf16_index[p] >= 0).
__global__ void __launch_bounds__(256) k_adam(float *p, const float *g, float *m, float *v, size_t n, float lr,
                                              float b1, float b2, float eps, float wd, float c1, float c2,
                                              const int *bf16_index, __bf16 *wb) {
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float upd = (mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i];
        const float pn = p[i] - lr * upd;
        p[i] = pn;
        const int bi = bf16_index[i];
        if (bi >= 0) wb[bi] = static_cast<__bf16>(pn);
    }
}

__global__ void __launch_bounds__(256) k_to_bf16(const float *p, const int *bf16_index, size_t n, __bf16 *wb) {
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256) {
        const int bi = bf16_index[i];
        if (bi >= 0) wb[bi] = static_cast<__bf16>(p[i]);
    }
}


// ---------------------------------------------------------------------------------------------
// Fused inference: one workgroup carries 64 rows through every layer. The activations stay in
// LDS (64 x 256 BF16, rows padded to 264 to spread the banks); each of the 4 wavefronts computes
// a quarter of the next layer's columns for all 64 rows (up to 2 x 2 MFMA tiles of 32 x 32,
// K in steps of 16 in increasing order, exactly as k_gemm accumulates), weight fragments come
// straight from global memory four k-steps ahead (the whole network is 338 KB: L2-resident), and
// the epilogue (bias,
// GELU, BF16) writes the next layer's input back into LDS. Same arithmetic as the layered forward,
// so the outputs are bit-identical to it.
// ---------------------------------------------------------------------------------------------
constexpr int kFuseRows = 64, kFuseTm = kFuseRows / 32, kFuseLd = 264;
constexpr int kFusePf = 4;  // k-steps of weight fragments in flight per wavefront

// One layer's K loop of a wavefront: kt = K / 16 k-steps, each one A fragment per row tile from
// LDS and the column tiles' weight fragments from global memory, loaded kFusePf steps ahead. With
// KC > 0 (the network's widths) the loop is unrolled whole, so no loop back-edge makes the
// compiler wait for every fragment in flight; KC = 0 takes any K (a multiple of 16, at most 256).
template <int KC>
__device__ __forceinline__ void fused_k(const __bf16 (&h)[kFuseRows][kFuseLd], const __bf16 *const (&wrow)[2], int tn_n,
                                        floatx16 (&acc)[kFuseTm][2], int kdyn = 0) {
    const int lane = threadIdx.x & 63;
    const int K = KC > 0 ? KC : kdyn;
    const int kt = K / 16;
    bf16x8 q[kFusePf][2];
#pragma unroll
    for (int s = 0; s < kFusePf; ++s)
        if (s < kt) {
            q[s][0] = *reinterpret_cast<const bf16x8 *>(wrow[0] + 16 * s);
            q[s][1] = tn_n > 1 ? *reinterpret_cast<const bf16x8 *>(wrow[1] + 16 * s) : q[s][0];
        }
    auto step = [&](int i, int s) {  // k-step i, its fragments in slot s
        const bf16x8 b0 = q[s][0], b1 = q[s][1];
        if (i + kFusePf < kt) {
            q[s][0] = *reinterpret_cast<const bf16x8 *>(wrow[0] + 16 * (i + kFusePf));
            if (tn_n > 1) q[s][1] = *reinterpret_cast<const bf16x8 *>(wrow[1] + 16 * (i + kFusePf));
        }
#pragma unroll
        for (int tm = 0; tm < kFuseTm; ++tm) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&h[32 * tm + (lane & 31)][16 * i + 8 * (lane >> 5)]);
            acc[tm][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b0, acc[tm][0], 0, 0, 0);
            if (tn_n > 1) acc[tm][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b1, acc[tm][1], 0, 0, 0);
        }
    };
    if constexpr (KC > 0) {
#pragma unroll
        for (int i = 0; i < KC / 16; ++i) step(i, i % kFusePf);
    } else {
        for (int i = 0; i < kt; i += kFusePf)
#pragma unroll
            for (int s = 0; s < kFusePf; ++s)
                if (i + s < kt) step(i + s, s);
    }
}

__global__ void __launch_bounds__(256) k_mlp_forward_fused(MlpFusedArgs g) {
    __shared__ __attribute__((aligned(16))) __bf16 h[kFuseRows][kFuseLd];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = blockIdx.x * kFuseRows;
    // input rows -> BF16, zero padded to the first layer's width
    const int k0w = g.kp[0];
    for (int e = threadIdx.x; e < kFuseRows * k0w; e += 256) {
        const int r = e / k0w, c = e - r * k0w;
        const int gr = r0 + r;
        const float v = (gr < g.M && c < g.n_in) ? g.x[static_cast<size_t>(gr) * g.n_in + c] : 0.0f;
        h[r][c] = static_cast<__bf16>(v);
    }
    __syncthreads();
    for (int l = 0; l < g.layers; ++l) {
        const int K = g.kp[l], N = g.n_out[l];
        const int ncw = N / 4;           // columns per wavefront: 64 or 32
        const int tn_n = ncw / 32;       // column tiles per wavefront: 2 or 1
        const int c0 = wave * ncw;
        const __bf16 *W = g.w[l];
        floatx16 acc[kFuseTm][2];
#pragma unroll
        for (int tm = 0; tm < kFuseTm; ++tm)
#pragma unroll
            for (int tn = 0; tn < 2; ++tn) acc[tm][tn] = floatx16{};
        // weight fragments of column tile tn, kFusePf k-steps ahead of their use
        const __bf16 *wrow[2] = {W + static_cast<size_t>(c0 + (lane & 31)) * K + 8 * (lane >> 5),
                                 W + static_cast<size_t>(c0 + 32 + (lane & 31)) * K + 8 * (lane >> 5)};
        if (K == 256) fused_k<256>(h, wrow, tn_n, acc);
        else if (K == 32) fused_k<32>(h, wrow, tn_n, acc);
        else fused_k<0>(h, wrow, tn_n, acc, K);
        __syncthreads();  // every wavefront is done reading this layer's input
        const bool last = l == g.layers - 1;
#pragma unroll
        for (int tn = 0; tn < 2; ++tn) {
            if (tn >= tn_n) continue;
            const int col = c0 + 32 * tn + (lane & 31);
            const float bj = g.bias[l][col];
#pragma unroll
            for (int tm = 0; tm < kFuseTm; ++tm)
#pragma unroll 4
                for (int r = 0; r < 16; ++r) {
                    const int row = 32 * tm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const float v = acc[tm][tn][r] + bj;
                    if (last) {
                        if (r0 + row < g.M) g.y[static_cast<size_t>(r0 + row) * N + col] = v;
                    } else {
                        h[row][col] = static_cast<__bf16>(gelu_f(v));
                    }
                }
        }
        __syncthreads();  // the next layer's input is complete
    }
}

}  // namespace

hipError_t launch_mlp_gemm(const MlpGemm &m, hipStream_t s) {
    GemmArgs g{};
    g.A = static_cast<const __bf16 *>(m.A);
    g.B = static_cast<const __bf16 *>(m.B);
    g.M = m.M;
    g.N = m.N;
    g.K = m.K;
    g.lda = m.lda;
    g.ldb = m.ldb;
    g.epi = m.epi;
    g.bias = m.bias;
    g.aux = m.aux;
    g.out_f32 = m.out_f32;
    g.out_bf16 = static_cast<__bf16 *>(m.out_bf16);
    g.ldc = m.ldc;
    if (m.M <= 0 || m.N <= 0) return hipSuccess;
    // 128 x 128 workgroup tiles (2 x 2 MFMA tiles per wavefront) when they still give every CU
    // several workgroups; otherwise 64 x 64 tiles balance better
    const long tiles128 = static_cast<long>((m.M + 127) / 128) * ((m.N + 127) / 128);
    if (m.M >= 128 && m.N >= 128 && tiles128 >= 2048) {
        dim3 grid((m.N + 127) / 128, (m.M + 127) / 128);
        if (!m.a_t && !m.b_t) k_gemm<false, false, 2><<<grid, 256, 0, s>>>(g);
        else if (!m.a_t && m.b_t) k_gemm<false, true, 2><<<grid, 256, 0, s>>>(g);
        else if (m.a_t && !m.b_t) k_gemm<true, false, 2><<<grid, 256, 0, s>>>(g);
        else k_gemm<true, true, 2><<<grid, 256, 0, s>>>(g);
        return hipGetLastError();
    }
    dim3 grid((m.N + 63) / 64, (m.M + 63) / 64);
    if (!m.a_t && !m.b_t) k_gemm<false, false, 1><<<grid, 256, 0, s>>>(g);
    else if (!m.a_t && m.b_t) k_gemm<false, true, 1><<<grid, 256, 0, s>>>(g);
    else if (m.a_t && !m.b_t) k_gemm<true, false, 1><<<grid, 256, 0, s>>>(g);
    else k_gemm<true, true, 1><<<grid, 256, 0, s>>>(g);
    return hipGetLastError();
}

hipError_t launch_mlp_mse(const float *y, const float *t, size_t n, float scale, float *g, void *gb,
                          double *partials, int blocks, hipStream_t s) {
    k_mse<<<blocks, 256, 0, s>>>(y, t, n, scale, g, static_cast<__bf16 *>(gb), partials);
    return hipGetLastError();
}

hipError_t launch_mlp_colsum(const float *g, int M, int N, float *db, hipStream_t s) {
    k_colsum<<<(N + 255) / 256, 256, 0, s>>>(g, M, N, db);
    return hipGetLastError();
}

hipError_t launch_mlp_gather(const float *x, int k_in, int ld, const float *t, int n_t, const int *idx, int rows,
                             void *xb, float *tb, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    k_gather<<<rows, 256, 0, s>>>(x, k_in, ld, t, n_t, idx, rows, static_cast<__bf16 *>(xb), tb);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                           float eps, float wd, float c1, float c2, const int *bf16_index, void *wb, hipStream_t s) {
    k_adam<<<512, 256, 0, s>>>(p, g, m, v, n, lr, b1, b2, eps, wd, c1, c2, bf16_index, static_cast<__bf16 *>(wb));
    return hipGetLastError();
}

hipError_t launch_mlp_to_bf16(const float *p, const int *bf16_index, size_t n, void *wb, hipStream_t s) {
    k_to_bf16<<<512, 256, 0, s>>>(p, bf16_index, n, static_cast<__bf16 *>(wb));
    return hipGetLastError();
}

int mlp_fused_supported(int layers, const int *kp, const int *n_out) {
    if (layers < 1 || layers > kMlpMaxLayers) return 0;
    for (int l = 0; l < layers; ++l) {
        if (kp[l] % 16 != 0 || kp[l] > 256) return 0;
        if (n_out[l] != 128 && n_out[l] != 256) return 0;
        if (l + 1 < layers && kp[l + 1] != n_out[l]) return 0;
    }
    return 1;
}

hipError_t launch_mlp_forward_fused(const MlpFusedArgs &a, hipStream_t s) {
    if (a.M <= 0) return hipSuccess;
    k_mlp_forward_fused<<<(a.M + kFuseRows - 1) / kFuseRows, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace emfd
