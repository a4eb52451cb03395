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

}  // namespace emfd
