// Far-field radiation integrals on the matrix cores.
//
// N(r) = sum_p J_p e^{i k r.(p - c)}, L(r) = sum_p M_p e^{i k r.(p - c)} over the Huygens box
// (J = w n x H, M = -w n x E from the stage-2 phasors), then |F|^2 = |(N - (N.r) r) - r x L|^2
// (shaders/ntff_pattern.comp:45-110 and ntff_reduce.comp:19-39 define the quantity). On a face
// with normal axis a and in-plane axes u, v the phase factors: e^{i k (r_a a0 + r_u u + r_v v)}, so
//
//   N_face(r) = e^{i k r_a a0} sum_u e^{i k r_u u} T(r, u),   T(r, u) = sum_v e^{i k r_v v} J(u, v),
//
// and T is a matrix product: rows (u, component, re / im) of the currents against a phase
// matrix of columns r and rows (v, cos / sin). It runs on v_mfma_f32_32x32x16_f16 with both
// operands split into an FP16 high and low part (a = a_hi + a_lo; the three products hi.hi,
// hi.lo, lo.hi are exact in FP32 and accumulate in FP32, about 22 significant bits). The
// currents are scaled by a power of two into [-1, 1] first, the phase factors are in [-1, 1].
// The u sum, the face phase and the 1 / scale run in FP32 in the epilogue; the six faces' N, L
// are added and |F|^2 formed by a second kernel. The radiated power, 1/2 sum w Re(E x H*).n,
// is a wavefront-reduced sum in FP64 partials.
#include "internal.h"

namespace emfd {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kDirTile = 256;  // directions per workgroup: 8 wavefronts of 32
constexpr int kRowsPerComp = 32;  // 16 u values x (re, im)

// Face point (iu, iv) of face f -> index into the phasor planes (the order ntff_point() walks:
// faces in turn, u fastest).
__device__ __forceinline__ int face_point(const FfFace &F, int iu, int iv) { return F.first + iv * F.U + iu; }

// Largest |phasor| over the stage-2 E and H planes (12 planes of n floats): one partial per block.
__global__ void __launch_bounds__(256) k_ff_absmax(const float *ph, size_t n, float *partials) {
    const size_t total = 12 * n;
    float m = 0.0f;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; i < total; i += static_cast<size_t>(gridDim.x) * 256)
        m = fmaxf(m, fabsf(ph[12 * n + i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    __shared__ float w[4];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = fmaxf(fmaxf(w[0], w[1]), fmaxf(w[2], w[3]));
}

// Power-of-two scale 2^-e with 2 max |phasor| 2^-e < 1 (a current is at most twice the largest
// phasor component: w <= 1, n a unit axis), from the absmax partials.
__device__ __forceinline__ float ff_scale(const float *partials, int n_partials) {
    float m = 0.0f;
    for (int i = 0; i < n_partials; ++i) m = fmaxf(m, partials[i]);
    if (!(m > 0.0f)) return 1.0f;
    int e;
    frexpf(m, &e);  // m < 2^e
    return ldexpf(1.0f, -(e + 1));
}

// Operands of every face: for point (iu, iv) and component c (0..2 J, 3..5 M) the current
// x + i y (times the scale) goes to row (iu, c, re) as (x, -y) and to row (iu, c, im) as (y, x)
// at k = 2 iv, 2 iv + 1, split into FP16 high and low parts. Layout (halves): per face, per block
// of 16 u values, per component, 32 rows (2 iu_local + part) of kp halves. Padding (u past U,
// k past 2V) is never written: the buffers are zeroed when they are allocated.
__global__ void __launch_bounds__(256) k_ff_operands(FfGeom g, const float *ph, int np, const float *absmax,
                                                     int n_absmax, _Float16 *a_hi, _Float16 *a_lo) {
    const int f = blockIdx.y;
    const FfFace &F = g.f[f];
    const int t = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (t >= F.U * F.V) return;
    const float s = ff_scale(absmax, n_absmax);
    const int iu = t % F.U, iv = t / F.U;
    const int i = face_point(F, iu, iv);
    const float w = ((iu == 0 || iu == F.U - 1) ? 0.5f : 1.0f) * ((iv == 0 || iv == F.V - 1) ? 0.5f : 1.0f);
    const float sgn = F.side ? 1.0f : -1.0f;  // outward normal +-e_a
    const size_t n = static_cast<size_t>(np);
    // q = 0: E re, 1: E im, 2: H re, 3: H im (stage-2 planes 12 + 3 q + c)
    float v[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) v[q][c] = ph[(12 + 3 * q + c) * n + i];
    // n x V for n = sgn e_a: component (a+1)%3 = -sgn V[(a+2)%3], (a+2)%3 = sgn V[(a+1)%3], a: 0
    const int a = F.ax, a1 = (a + 1) % 3, a2 = (a + 2) % 3;
    float cur[6][2];  // [c][re, im]
#pragma unroll
    for (int c = 0; c < 6; ++c) cur[c][0] = cur[c][1] = 0.0f;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        const float *H = v[2 + part], *Ev = v[part];
        const float hx[3] = {H[0], H[1], H[2]}, ex[3] = {Ev[0], Ev[1], Ev[2]};
        // J = w n x H, M = -w n x E
        float jv[3] = {0.0f, 0.0f, 0.0f}, mv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float nh = c == a1 ? -sgn * hx[a2] : (c == a2 ? sgn * hx[a1] : 0.0f);
            const float ne = c == a1 ? -sgn * ex[a2] : (c == a2 ? sgn * ex[a1] : 0.0f);
            jv[c] = w * nh;
            mv[c] = -w * ne;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cur[c][part] = jv[c] * s;
            cur[3 + c][part] = mv[c] * s;
        }
    }
    const int ub = iu >> 4, row_u = 2 * (iu & 15);
    const size_t base = F.a_off + static_cast<size_t>(ub) * 6 * kRowsPerComp * F.kp;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        const float x = cur[c][0], y = cur[c][1];
        // (row re: x at 2 iv, -y at 2 iv + 1), (row im: y, x)
        const float vals[2][2] = {{x, -y}, {y, x}};
#pragma unroll
        for (int part = 0; part < 2; ++part) {
            const size_t r = base + (static_cast<size_t>(c) * kRowsPerComp + row_u + part) * F.kp + 2 * iv;
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                const float val = vals[part][tt];
                const _Float16 hi = static_cast<_Float16>(val);
                const _Float16 lo = static_cast<_Float16>(val - static_cast<float>(hi));
                a_hi[r + tt] = hi;
                a_lo[r + tt] = lo;
            }
        }
    }
}

// One workgroup: 256 directions (8 wavefronts x 32) of one face, every block of 16 u values in
// turn, K in steps of 16 (8 (v, cos / sin) pairs per lane and step). Per step the 6 x 32 x 16
// operand rows (hi and lo, 12 KB) are staged in LDS, shared by the 8 wavefronts; each wavefront
// forms its 32 directions' phase column block in registers. MFMA layout (32x32x16): lane l holds
// rows l % 32 of A and columns l % 32 of B at k = 8 (l / 32) + e, e = 0..7, and receives
// D[8 (r / 4) + 4 (l / 32) + r % 4][l % 32] in accumulator r.
__global__ void __launch_bounds__(512) k_ff_face(FfGeom g, const float *dirs, int n_dirs, float kw,
                                                 const float *absmax, int n_absmax, const _Float16 *a_hi,
                                                 const _Float16 *a_lo, float *partials) {
    __shared__ __attribute__((aligned(16))) _Float16 sA[2][6][kRowsPerComp][16];  // [hi/lo][c][row][k]
    const int f = blockIdx.y;
    const FfFace &F = g.f[f];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int d = static_cast<int>(blockIdx.x) * kDirTile + 32 * wave + (lane & 31);
    const bool dok = d < n_dirs;
    const float ru = dok ? dirs[4 * d + F.u] : 0.0f, rv = dok ? dirs[4 * d + F.v] : 0.0f,
                ra = dok ? dirs[4 * d + F.ax] : 0.0f;
    const float inv_s = 1.0f / ff_scale(absmax, n_absmax);
    // phase step along v for this lane's direction
    float cd, sd;
    sincosf(kw * rv, &sd, &cd);
    const int half = lane >> 5;  // k = 8 half + e within a step
    float acc_out[6][2];
#pragma unroll
    for (int c = 0; c < 6; ++c) acc_out[c][0] = acc_out[c][1] = 0.0f;
    const int nks = F.kp / 16;
    const int tid = threadIdx.x;
    for (int ub = 0; ub < F.ublocks; ++ub) {
        const _Float16 *blk_hi = a_hi + F.a_off + static_cast<size_t>(ub) * 6 * kRowsPerComp * F.kp;
        const _Float16 *blk_lo = a_lo + F.a_off + static_cast<size_t>(ub) * 6 * kRowsPerComp * F.kp;
        f32x16 acc[6];
#pragma unroll
        for (int c = 0; c < 6; ++c)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[c][r] = 0.0f;
        // staging: 768 chunks of 16 B per step (hi/lo x 6 comps x 32 rows x 2 halves of 8 k)
        auto chunk_src = [&](int q, int ks) {
            const int hl = q / 384, rem = q % 384;
            const int c = rem >> 6, row = (rem & 63) >> 1, h8 = rem & 1;
            const _Float16 *src = (hl ? blk_lo : blk_hi) + (static_cast<size_t>(c) * kRowsPerComp + row) * F.kp +
                                  ks * 16 + 8 * h8;
            return src;
        };
        auto chunk_dst = [&](int q) {
            const int hl = q / 384, rem = q % 384;
            const int c = rem >> 6, row = (rem & 63) >> 1, h8 = rem & 1;
            return &sA[hl][c][row][8 * h8];
        };
        uint4 pre0 = *reinterpret_cast<const uint4 *>(chunk_src(tid, 0));
        uint4 pre1 = tid < 256 ? *reinterpret_cast<const uint4 *>(chunk_src(tid + 512, 0)) : make_uint4(0, 0, 0, 0);
        for (int ks = 0; ks < nks; ++ks) {
            __syncthreads();  // every wavefront is done with the previous step's operands
            *reinterpret_cast<uint4 *>(chunk_dst(tid)) = pre0;
            if (tid < 256) *reinterpret_cast<uint4 *>(chunk_dst(tid + 512)) = pre1;
            __syncthreads();
            if (ks + 1 < nks) {
                pre0 = *reinterpret_cast<const uint4 *>(chunk_src(tid, ks + 1));
                if (tid < 256) pre1 = *reinterpret_cast<const uint4 *>(chunk_src(tid + 512, ks + 1));
            }
            // phase column block: k = 16 ks + 8 half + e -> iv = 8 ks + 4 half + e / 2, cos (e even) / sin
            const int iv0 = 8 * ks + 4 * half;
            float c0, s0;
            sincosf(kw * rv * (F.v0 + static_cast<float>(iv0)), &s0, &c0);
            f16x8 b_hi, b_lo;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const _Float16 ch = static_cast<_Float16>(c0), sh = static_cast<_Float16>(s0);
                b_hi[2 * e] = ch;
                b_hi[2 * e + 1] = sh;
                b_lo[2 * e] = static_cast<_Float16>(c0 - static_cast<float>(ch));
                b_lo[2 * e + 1] = static_cast<_Float16>(s0 - static_cast<float>(sh));
                const float cn = c0 * cd - s0 * sd, sn = s0 * cd + c0 * sd;
                c0 = cn;
                s0 = sn;
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const f16x8 ah = *reinterpret_cast<const f16x8 *>(&sA[0][c][lane & 31][8 * half]);
                const f16x8 al = *reinterpret_cast<const f16x8 *>(&sA[1][c][lane & 31][8 * half]);
                acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, b_hi, acc[c], 0, 0, 0);
                acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, b_lo, acc[c], 0, 0, 0);
                acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, b_hi, acc[c], 0, 0, 0);
            }
        }
        // epilogue: rows 8 (r / 4) + 4 half + r % 4 = 2 iu_local + part
#pragma unroll
        for (int r = 0; r < 16; r += 2) {
            const int iu = 16 * ub + (8 * (r >> 2) + 4 * half + (r & 3)) / 2;
            float cu, su;
            sincosf(kw * ru * (F.u0 + static_cast<float>(iu)), &su, &cu);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const float tr = acc[c][r], ti = acc[c][r + 1];
                acc_out[c][0] += tr * cu - ti * su;
                acc_out[c][1] += tr * su + ti * cu;
            }
        }
    }
    // lanes l and l + 32 hold the same direction and complementary u values
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        acc_out[c][0] += __shfl_xor(acc_out[c][0], 32);
        acc_out[c][1] += __shfl_xor(acc_out[c][1], 32);
    }
    if (half == 0 && dok) {
        float ca, sa;
        sincosf(kw * ra * F.a0, &sa, &ca);
        float *o = partials + (static_cast<size_t>(f) * n_dirs + d) * 12;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const float xr = acc_out[c][0] * inv_s, xi = acc_out[c][1] * inv_s;
            const float yr = xr * ca - xi * sa, yi = xr * sa + xi * ca;
            // layout Nr[3] Ni[3] Lr[3] Li[3]
            const int base = c < 3 ? 0 : 6, comp = c % 3;
            o[base + comp] = yr;
            o[base + 3 + comp] = yi;
        }
    }
}

// |F|^2 = |(N - (N.r) r) - r x L|^2 from the six faces' N, L.
__global__ void __launch_bounds__(256) k_ff_intensity(const float *dirs, int n_dirs, const float *partials,
                                                      float *out) {
    const int d = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (d >= n_dirs) return;
    float s[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) s[j] = 0.0f;
    for (int f = 0; f < 6; ++f) {
        const float *p = partials + (static_cast<size_t>(f) * n_dirs + d) * 12;
#pragma unroll
        for (int j = 0; j < 12; ++j) s[j] += p[j];
    }
    const float r[3] = {dirs[4 * d], dirs[4 * d + 1], dirs[4 * d + 2]};
    float sum = 0.0f;
#pragma unroll
    for (int part = 0; part < 2; ++part) {
        const float *N = s + 3 * part, *L = s + 6 + 3 * part;
        const float nr = N[0] * r[0] + N[1] * r[1] + N[2] * r[2];
        const float rxl[3] = {r[1] * L[2] - r[2] * L[1], r[2] * L[0] - r[0] * L[2], r[0] * L[1] - r[1] * L[0]};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const float fj = (N[j] - nr * r[j]) - rxl[j];
            sum += fj * fj;
        }
    }
    out[d] = sum;
}

// 1/2 sum w Re(E x H*).n over the box: per point in FP32, per wavefront summed by butterfly
// shuffles, one FP64 partial per workgroup.
__global__ void __launch_bounds__(256) k_ff_power(NtffBox box, const float *ph, double *partials) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    float p = 0.0f;
    if (i < box.n_points) {
        int f = 0;
#pragma unroll
        for (int q = 1; q < 6; ++q)
            if (i >= box.face_offset[q]) f = q;
        const int a = f >> 1, side = f & 1;
        const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        const int U = box.bmax[u] - box.bmin[u] + 1, V = box.bmax[v] - box.bmin[v] + 1;
        const int local = i - box.face_offset[f], iu = local % U, iv = local / U;
        const float w = ((iu == 0 || iu == U - 1) ? 0.5f : 1.0f) * ((iv == 0 || iv == V - 1) ? 0.5f : 1.0f);
        const size_t n = static_cast<size_t>(box.n_points);
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        // (E x H*)_a = E_a1 conj(H_a2) - E_a2 conj(H_a1); its real part
        const float er1 = ph[(12 + a1) * n + i], ei1 = ph[(15 + a1) * n + i];
        const float er2 = ph[(12 + a2) * n + i], ei2 = ph[(15 + a2) * n + i];
        const float hr1 = ph[(18 + a1) * n + i], hi1 = ph[(21 + a1) * n + i];
        const float hr2 = ph[(18 + a2) * n + i], hi2 = ph[(21 + a2) * n + i];
        const float re = (er1 * hr2 + ei1 * hi2) - (er2 * hr1 + ei2 * hi1);
        p = (side ? w : -w) * re;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o);
    __shared__ float wsum[4];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = p;
    __syncthreads();
    if (threadIdx.x == 0)
        partials[blockIdx.x] = (static_cast<double>(wsum[0]) + wsum[1]) + (static_cast<double>(wsum[2]) + wsum[3]);
}

}  // namespace

FfGeom ff_geometry(const NtffBox &box, size_t *a_halves) {
    FfGeom g{};
    size_t off = 0;
    const float c[3] = {0.5f * static_cast<float>(box.bmin[0] + box.bmax[0]),
                        0.5f * static_cast<float>(box.bmin[1] + box.bmax[1]),
                        0.5f * static_cast<float>(box.bmin[2] + box.bmax[2])};
    for (int f = 0; f < 6; ++f) {
        FfFace &F = g.f[f];
        F.ax = f >> 1;
        F.side = f & 1;
        F.u = F.ax == 0 ? 1 : 0;
        F.v = F.ax == 2 ? 1 : 2;
        F.U = box.bmax[F.u] - box.bmin[F.u] + 1;
        F.V = box.bmax[F.v] - box.bmin[F.v] + 1;
        F.first = box.face_offset[f];
        F.ublocks = (F.U + 15) / 16;
        F.kp = (2 * F.V + 15) / 16 * 16;
        F.a0 = static_cast<float>(F.side ? box.bmax[F.ax] : box.bmin[F.ax]) - c[F.ax];
        F.u0 = static_cast<float>(box.bmin[F.u]) - c[F.u];
        F.v0 = static_cast<float>(box.bmin[F.v]) - c[F.v];
        F.a_off = off;
        off += static_cast<size_t>(F.ublocks) * 6 * kRowsPerComp * F.kp;
    }
    if (a_halves) *a_halves = off;
    return g;
}

int ff_absmax_blocks(int n_points) {
    const int b = (12 * n_points + 256 * 64 - 1) / (256 * 64);
    return b < 1 ? 1 : (b > 512 ? 512 : b);
}

hipError_t launch_far_field(const NtffBox &box, const FfGeom &g, const float *phasors, const float *dirs, int n_dirs,
                            float k, float *absmax, _Float16 *a_hi, _Float16 *a_lo, float *face_partials,
                            float *out_f2, double *power_partials, hipStream_t s) {
    const int np = box.n_points;
    const int nab = ff_absmax_blocks(np);
    if (n_dirs > 0 && np > 0) {
        k_ff_absmax<<<nab, 256, 0, s>>>(phasors, static_cast<size_t>(np), absmax);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        int most = 0;
        for (int f = 0; f < 6; ++f) most = most > g.f[f].U * g.f[f].V ? most : g.f[f].U * g.f[f].V;
        k_ff_operands<<<dim3((most + 255) / 256, 6), 256, 0, s>>>(g, phasors, np, absmax, nab, a_hi, a_lo);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        k_ff_face<<<dim3((n_dirs + kDirTile - 1) / kDirTile, 6), 512, 0, s>>>(g, dirs, n_dirs, k, absmax, nab, a_hi,
                                                                              a_lo, face_partials);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        k_ff_intensity<<<(n_dirs + 255) / 256, 256, 0, s>>>(dirs, n_dirs, face_partials, out_f2);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (power_partials && np > 0) {
        k_ff_power<<<(np + 255) / 256, 256, 0, s>>>(box, phasors, power_partials);
        return hipGetLastError();
    }
    return hipSuccess;
}

}  // namespace emfd
