// Convolutional PML (opt-in, EMFD_ABSORBER_CPML): the recursive-convolution stretched-coordinate
// absorber of Roden and Gedney (2000) with the complex-frequency-shifted stretch
// s = 1 + sigma / (alpha + j w) (kappa = 1), in the slabs the graded absorber would occupy. The
// Yee passes then run with the absorber's conductivity switched off (vacuum coefficients); after
// each pass these kernels update the auxiliary fields of every axis and add them to the field:
//
//   after Faraday:  psi_Bv = b psi_Bv + a d_u E_w;  B_v += dt psi_Bv
//                   psi_Bw = b psi_Bw + a d_u E_v;  B_w -= dt psi_Bw      (forward differences)
//   after Ampere:   psi_Ev = b psi_Ev + a d_u B_w;  E_v -= Cb_v psi_Ev
//                   psi_Ew = b psi_Ew + a d_u B_v;  E_w += Cb_w psi_Ew    (backward differences)
//
// with v = u + 1, w = u + 2 (mod 3): the terms of the curl that differentiate along u. The axes
// run one launch each, x then y then z, so that a corner texel's corrections are added in a fixed
// order. One thread per slab texel, x fastest; the slab of axis u is the planes 0 .. L-1 and
// n-1-L .. n-1 along u (2 L + 1 planes).
#include "yee_device.h"

namespace emfd {
namespace {

template <int U, bool EPH>
__global__ void __launch_bounds__(256) k_cpml(LatticeArgs a, FieldPtrs f, const float *__restrict__ coef,
                                              float *__restrict__ psi, int stride) {
    constexpr int V = (U + 1) % 3, W = (U + 2) % 3;
    const int L = a.pml;
    int n[3] = {a.nx, a.ny, a.nz};
    int dims[3] = {a.nx, a.ny, a.nz};
    dims[U] = 2 * L + 1;
    const int cells = dims[0] * dims[1] * dims[2];
    const int t = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (t >= cells) return;
    int q[3] = {t % dims[0], (t / dims[0]) % dims[1], t / (dims[0] * dims[1])};
    q[U] = q[U] < L ? q[U] : n[U] - 1 - 2 * L + q[U];
    const size_t sy = static_cast<size_t>(a.nx), sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = static_cast<size_t>(q[2]) * sz + static_cast<size_t>(q[1]) * sy + static_cast<size_t>(q[0]);
    const size_t su = U == 0 ? 1 : (U == 1 ? sy : sz);
    const float *c = coef + (static_cast<size_t>(U * 2 + (EPH ? 0 : 1)) * stride + q[U]) * 2;
    const float bq = c[0], aq = c[1];  // b, a of this position
    float *pv = psi + (EPH ? 2 : 0) * static_cast<size_t>(cells) + t;
    float *pw = psi + (EPH ? 3 : 1) * static_cast<size_t>(cells) + t;
    const float dt = a.dt;
    if (!EPH) {
        const bool in = q[U] + 1 < n[U];
        const float ew0 = f.e[W][p], ev0 = f.e[V][p];
        const float ew1 = in ? f.e[W][p + su] : 0.0f;
        const float ev1 = in ? f.e[V][p + su] : 0.0f;
        const float nv = bq * *pv + aq * (ew1 - ew0);
        const float nw = bq * *pw + aq * (ev1 - ev0);
        *pv = nv;
        *pw = nw;
        f.b[V][p] = f.b[V][p] + dt * nv;
        f.b[W][p] = f.b[W][p] - dt * nw;
    } else {
        const bool in = q[U] > 0;
        const float bw0 = f.b[W][p], bv0 = f.b[V][p];
        const float bw1 = in ? f.b[W][p - su] : 0.0f;
        const float bv1 = in ? f.b[V][p - su] : 0.0f;
        const float nv = bq * *pv + aq * (bw0 - bw1);
        const float nw = bq * *pw + aq * (bv0 - bv1);
        *pv = nv;
        *pw = nw;
        float ca[3], cb[3];
        edge_coefficients(a, q[0], q[1], q[2], a.mat_id ? load_id(a, p) : 0u, ca, cb);
        f.e[V][p] = f.e[V][p] - cb[V] * nv;
        f.e[W][p] = f.e[W][p] + cb[W] * nw;
    }
}

// ---------------------------------------------------------------------------------------------
// The same absorber around the one-pass march, whose update (and sparse passes) treat every texel
// as vacuum or material. src holds the fields of step n, dst the march's output. Per texel, every
// slab's correction is applied in the order x, y, z with the operations of k_cpml, so the fields
// are those of the two passes with the per-axis launches, bit for bit:
//   k_cpml_march_b: the slabs' union; psi_B from E^n (src), B corrections into dst;
//   k_cpml_march_e: the union widened by one plane inward on the low sides (every texel whose
//     curl reads a corrected B): E^{n+1} recomputed from the corrected B as the Ampere pass
//     computes it, then the psi_E corrections.
// ---------------------------------------------------------------------------------------------

// Texel t of the shell of lo planes at the low and hi planes at the high end of every axis: the
// x shell (x fastest), then the y shell without it, then the z shell without both. False past it.
__device__ __forceinline__ bool shell_texel(const LatticeArgs &a, int lo, int hi, long t, int (&q)[3]) {
    const int nx = a.nx, ny = a.ny, nz = a.nz, w = lo + hi;
    auto sh = [&](long s, int n) { return static_cast<int>(s < lo ? s : n - w + s); };
    const long p0 = static_cast<long>(w) * ny * nz;
    if (t < p0) {
        q[0] = sh(t % w, nx);
        q[1] = static_cast<int>((t / w) % ny);
        q[2] = static_cast<int>(t / (static_cast<long>(w) * ny));
        return true;
    }
    t -= p0;
    const int ix = nx - w, iy = ny - w;
    const long p1 = static_cast<long>(ix) * w * nz;
    if (t < p1) {
        q[0] = lo + static_cast<int>(t % ix);
        q[1] = sh((t / ix) % w, ny);
        q[2] = static_cast<int>(t / (static_cast<long>(ix) * w));
        return true;
    }
    t -= p1;
    if (t < static_cast<long>(ix) * iy * w) {
        q[0] = lo + static_cast<int>(t % ix);
        q[1] = lo + static_cast<int>((t / ix) % iy);
        q[2] = sh(t / (static_cast<long>(ix) * iy), nz);
        return true;
    }
    return false;
}

__host__ __device__ inline long shell_count(int nx, int ny, int nz, int lo, int hi) {
    const int w = lo + hi;
    return static_cast<long>(w) * ny * nz + static_cast<long>(nx - w) * w * nz + static_cast<long>(nx - w) * (ny - w) * w;
}

// The slab of axis U (planes 0 .. L-1 and n-1-L .. n-1) holds q: its auxiliary-field index (the
// layout of k_cpml's thread index) and base.
template <int U>
__device__ __forceinline__ bool slab_index(const LatticeArgs &a, const int (&q)[3], float *psi, float **base, long *idx) {
    const int L = a.pml, n[3] = {a.nx, a.ny, a.nz};
    if (q[U] >= L && q[U] < n[U] - 1 - L) return false;
    const long S = 2 * L + 1;
    const long per0 = S * n[1] * n[2], per1 = static_cast<long>(n[0]) * S * n[2];
    const long per = U == 0 ? per0 : (U == 1 ? per1 : static_cast<long>(n[0]) * n[1] * S);
    *base = psi + (U == 0 ? 0 : (U == 1 ? 4 * per0 : 4 * (per0 + per1)));
    int c[3] = {q[0], q[1], q[2]};
    long d[3] = {n[0], n[1], n[2]};
    c[U] = q[U] < L ? q[U] : q[U] - (n[U] - 1 - 2 * L);
    d[U] = S;
    *idx = c[0] + d[0] * (c[1] + d[1] * static_cast<long>(c[2]));
    (void)per;
    return true;
}

template <int U>
__device__ __forceinline__ long slab_cells(const LatticeArgs &a) {
    const long S = 2 * a.pml + 1;
    return U == 0 ? S * a.ny * a.nz : (U == 1 ? static_cast<long>(a.nx) * S * a.nz : static_cast<long>(a.nx) * a.ny * S);
}

template <int U>
__device__ __forceinline__ void march_b_axis(const LatticeArgs &a, const FieldPtrs &src, const float *coef, float *psi,
                                             int stride, const int (&q)[3], size_t p, float (&b)[3]) {
    constexpr int V = (U + 1) % 3, W = (U + 2) % 3;
    float *base;
    long t;
    if (!slab_index<U>(a, q, psi, &base, &t)) return;
    const int n[3] = {a.nx, a.ny, a.nz};
    const size_t su = U == 0 ? 1 : (U == 1 ? static_cast<size_t>(a.nx) : static_cast<size_t>(a.nx) * a.ny);
    const long cells = slab_cells<U>(a);
    const float *c = coef + (static_cast<size_t>(U * 2 + 1) * stride + q[U]) * 2;
    const float bq = c[0], aq = c[1];
    float *pv = base + t, *pw = base + cells + t;
    const bool in = q[U] + 1 < n[U];
    const float ew0 = src.e[W][p], ev0 = src.e[V][p];
    const float ew1 = in ? src.e[W][p + su] : 0.0f;
    const float ev1 = in ? src.e[V][p + su] : 0.0f;
    const float nv = bq * *pv + aq * (ew1 - ew0);
    const float nw = bq * *pw + aq * (ev1 - ev0);
    *pv = nv;
    *pw = nw;
    b[V] = b[V] + a.dt * nv;
    b[W] = b[W] - a.dt * nw;
}

template <int U>
__device__ __forceinline__ void march_e_axis(const LatticeArgs &a, const FieldPtrs &dst, const float *coef, float *psi,
                                             int stride, const int (&q)[3], size_t p, const float (&cb)[3],
                                             float (&e)[3]) {
    constexpr int V = (U + 1) % 3, W = (U + 2) % 3;
    float *base;
    long t;
    if (!slab_index<U>(a, q, psi, &base, &t)) return;
    const size_t su = U == 0 ? 1 : (U == 1 ? static_cast<size_t>(a.nx) : static_cast<size_t>(a.nx) * a.ny);
    const long cells = slab_cells<U>(a);
    const float *c = coef + (static_cast<size_t>(U * 2) * stride + q[U]) * 2;
    const float bq = c[0], aq = c[1];
    float *pv = base + 2 * cells + t, *pw = base + 3 * cells + t;
    const bool in = q[U] > 0;
    const float bw0 = dst.b[W][p], bv0 = dst.b[V][p];
    const float bw1 = in ? dst.b[W][p - su] : 0.0f;
    const float bv1 = in ? dst.b[V][p - su] : 0.0f;
    const float nv = bq * *pv + aq * (bw0 - bw1);
    const float nw = bq * *pw + aq * (bv0 - bv1);
    *pv = nv;
    *pw = nw;
    e[V] = e[V] - cb[V] * nv;
    e[W] = e[W] + cb[W] * nw;
}

// the far-field side columns keep what the planes hold (comp0: 0 for E, 3 for B)
__device__ __forceinline__ void cpml_side(const LatticeArgs &a, int comp0, const int (&q)[3], const float (&v)[3]) {
    if (!a.side.buf) return;
    for (int face = 0; face < 2; ++face)
        for (int c = 0; c < 3; ++c) {
            const long k = side_index(a.side, face, comp0 + c, q[0], q[1], q[2]);
            if (k >= 0) a.side.buf[k] = v[c];
        }
}

__global__ void __launch_bounds__(256) k_cpml_march_b(LatticeArgs a, FieldPtrs src, FieldPtrs dst,
                                                      const float *__restrict__ coef, float *__restrict__ psi, int stride) {
    int q[3];
    if (!shell_texel(a, a.pml, a.pml + 1, static_cast<long>(blockIdx.x) * 256 + threadIdx.x, q)) return;
    const size_t p = (static_cast<size_t>(q[2]) * a.ny + q[1]) * a.nx + q[0];
    // the components some slab of this texel corrects (slab u: components u + 1, u + 2)
    const int L = a.pml, n[3] = {a.nx, a.ny, a.nz};
    bool in[3], touch[3];
#pragma unroll
    for (int u = 0; u < 3; ++u) in[u] = q[u] < L || q[u] >= n[u] - 1 - L;
#pragma unroll
    for (int c = 0; c < 3; ++c) touch[c] = in[(c + 1) % 3] || in[(c + 2) % 3];
    float b[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) b[c] = touch[c] ? dst.b[c][p] : 0.0f;
    march_b_axis<0>(a, src, coef, psi, stride, q, p, b);
    march_b_axis<1>(a, src, coef, psi, stride, q, p, b);
    march_b_axis<2>(a, src, coef, psi, stride, q, p, b);
#pragma unroll
    for (int c = 0; c < 3; ++c)
        if (touch[c]) dst.b[c][p] = b[c];
    if (a.side.buf) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            if (!touch[c]) b[c] = dst.b[c][p];
        cpml_side(a, 3, q, b);
    }
}

__global__ void __launch_bounds__(256) k_cpml_march_e(LatticeArgs a, FieldPtrs src, FieldPtrs dst,
                                                      const float *__restrict__ coef, float *__restrict__ psi, int stride) {
    int q[3];
    if (!shell_texel(a, a.pml + 1, a.pml + 1, static_cast<long>(blockIdx.x) * 256 + threadIdx.x, q)) return;
    const int x = q[0], y = q[1], z = q[2];
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x;
    // E^{n+1} from the corrected B^{n+1/2} (the Ampere pass's operations, update_E.comp:35-40)
    float b0[3], bX[3], bY[3], bZ[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b0[c] = dst.b[c][p];
        bX[c] = x > 0 ? dst.b[c][p - 1] : 0.0f;
        bY[c] = y > 0 ? dst.b[c][p - sy] : 0.0f;
        bZ[c] = z > 0 ? dst.b[c][p - sz] : 0.0f;
    }
    const float curl[3] = {(b0[2] - bY[2]) - (b0[1] - bZ[1]), (b0[0] - bZ[0]) - (b0[2] - bX[2]),
                           (b0[1] - bX[1]) - (b0[0] - bY[0])};
    float ca[3], cb[3];
    edge_coefficients(a, x, y, z, a.mat_id ? load_id(a, p) : 0u, ca, cb);
    float e[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c] = ca[c] * src.e[c][p] + cb[c] * curl[c];
    march_e_axis<0>(a, dst, coef, psi, stride, q, p, cb, e);
    march_e_axis<1>(a, dst, coef, psi, stride, q, p, cb, e);
    march_e_axis<2>(a, dst, coef, psi, stride, q, p, cb, e);
#pragma unroll
    for (int c = 0; c < 3; ++c) dst.e[c][p] = e[c];
    cpml_side(a, 0, q, e);
}

template <bool EPH>
hipError_t launch_phase(const LatticeArgs &a, const FieldPtrs &f, const float *coef, float *psi, hipStream_t s) {
    const int stride = a.prof_stride;
    const size_t per[3] = {static_cast<size_t>(2 * a.pml + 1) * a.ny * a.nz,
                           static_cast<size_t>(a.nx) * (2 * a.pml + 1) * a.nz,
                           static_cast<size_t>(a.nx) * a.ny * (2 * a.pml + 1)};
    float *p0 = psi, *p1 = p0 + 4 * per[0], *p2 = p1 + 4 * per[1];
    k_cpml<0, EPH><<<static_cast<unsigned>((per[0] + 255) / 256), 256, 0, s>>>(a, f, coef, p0, stride);
    k_cpml<1, EPH><<<static_cast<unsigned>((per[1] + 255) / 256), 256, 0, s>>>(a, f, coef, p1, stride);
    k_cpml<2, EPH><<<static_cast<unsigned>((per[2] + 255) / 256), 256, 0, s>>>(a, f, coef, p2, stride);
    return hipGetLastError();
}

}  // namespace

size_t cpml_psi_floats(int nx, int ny, int nz, int pml) {
    const size_t S = static_cast<size_t>(2 * pml + 1);
    return 4 * (S * ny * nz + nx * S * nz + static_cast<size_t>(nx) * ny * S);
}

hipError_t launch_cpml(const LatticeArgs &a, const FieldPtrs &f, const float *coef, float *psi, bool e_phase,
                       hipStream_t s) {
    if (a.pml <= 0) return hipSuccess;
    return e_phase ? launch_phase<true>(a, f, coef, psi, s) : launch_phase<false>(a, f, coef, psi, s);
}

int cpml_march_supported(const LatticeArgs &a) {
    // the widened shell (configure's smallest lattice, 2 pml + 8 cells, always leaves room)
    const int w = 2 * a.pml + 2;
    return a.nx > w && a.ny > w && a.nz > w ? 1 : 0;
}

hipError_t launch_cpml_march(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, const float *coef,
                             float *psi, bool e_phase, hipStream_t s) {
    if (a.pml <= 0) return hipSuccess;
    if (!cpml_march_supported(a)) return hipErrorInvalidValue;
    const int stride = a.prof_stride;
    if (!e_phase) {
        const long n = shell_count(a.nx, a.ny, a.nz, a.pml, a.pml + 1);
        k_cpml_march_b<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(a, src, dst, coef, psi, stride);
    } else {
        const long n = shell_count(a.nx, a.ny, a.nz, a.pml + 1, a.pml + 1);
        k_cpml_march_e<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(a, src, dst, coef, psi, stride);
    }
    return hipGetLastError();
}

}  // namespace emfd
