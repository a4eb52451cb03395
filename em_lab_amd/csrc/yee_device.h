// Device helpers shared by the Yee kernels: absorber profile, phi(x), coefficient classes, the
// per-texel Faraday update and the Ampere coefficients. Each follows the reference shader it cites
// operation by operation; compiled without contraction, with det_expf for every exponential.
#pragma once

#include "expf.h"
#include "internal.h"

namespace emfd {

__device__ __forceinline__ float phi_f(float x) {
    // phi(x) = (1 - e^{-x}) / x, Taylor series below 1e-2 (shaders/common.glsl:50-54)
    const float series = 1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f)));
    const float exact = (1.0f - det_expf(-x)) / fmaxf(x, 1e-12f);
    return x >= 1e-2f ? exact : series;
}

__device__ __forceinline__ bool in_pml(const LatticeArgs &a, int x, int y, int z) {
    // shaders/common.glsl:40-43
    const int L = a.pml;
    return x < L || y < L || z < L || x >= a.nx - (L + 1) || y >= a.ny - (L + 1) || z >= a.nz - (L + 1);
}

// Absorber sigma at a staggered point: (profile_x + profile_y) + profile_z as pmlSigma sums it
// (shaders/common.glsl:31-36). hx/hy/hz select the +1/2 samples.
__device__ __forceinline__ float pml_sigma(const LatticeArgs &a, int x, int hx, int y, int hy, int z, int hz) {
    const int s = a.prof_stride;
    return a.prof[(2 * 0 + hx) * s + x] + a.prof[(2 * 1 + hy) * s + y] + a.prof[(2 * 2 + hz) * s + z];
}

__host__ __device__ __forceinline__ int pml_class(int i, int n, int L) {
    return i < L ? 1 + i : (i >= n - 1 - L ? 1 + L + (i - (n - 1 - L)) : 0);
}

__device__ __forceinline__ const float *pml_entry(const LatticeArgs &a, int x, int y, int z) {
    const int K = a.pml_k;
    const int e = pml_class(x, a.nx, a.pml) + K * (pml_class(y, a.ny, a.pml) + K * pml_class(z, a.nz, a.pml));
    return a.pml_tab + static_cast<size_t>(e) * kPmlStride;
}

__device__ __forceinline__ uint32_t load_id(const LatticeArgs &a, size_t i) {
    if (a.id_bytes == 2) return static_cast<const uint16_t *>(a.mat_id)[i];
    return static_cast<const uint32_t *>(a.mat_id)[i];
}

// Faraday update of one texel (shaders/update_B.comp:35-69). e0: E at p; exn/eyn/ezn: E at
// p+x, p+y, p+z (0 beyond the upper faces). id: class of p (0 when no table).
__device__ __forceinline__ void faraday(const LatticeArgs &a, int x, int y, int z, const float e0[3],
                                        const float exn[3], const float eyn[3], const float ezn[3], uint32_t id,
                                        float b[3]) {
    float t0x = eyn[2] - e0[2], t0y = ezn[0] - e0[0], t0z = exn[1] - e0[1];
    float t1x = ezn[1] - e0[1], t1y = exn[2] - e0[2], t1z = eyn[0] - e0[0];
    if (a.has_wire && id != 0u) {  // class 0 is vacuum: factors 1
        const float *w = a.table + static_cast<size_t>(id) * kTableStride + kWire;
        t0x *= w[0];
        t1x *= w[1];
        t0y *= w[2];
        t1y *= w[3];
        t0z *= w[4];
        t1z *= w[5];
    }
    const float cx = t0x - t1x, cy = t0y - t1y, cz = t0z - t1z;
    const float dt = a.dt;
    if (in_pml(a, x, y, z)) {
        // exp(-s dt) and dt phi(s dt), s = sigma_m at the B stagger, tabulated per depth class
        const float *t = pml_entry(a, x, y, z);
        b[0] = t[kPmlBa + 0] * b[0] - t[kPmlBb + 0] * cx;
        b[1] = t[kPmlBa + 1] * b[1] - t[kPmlBb + 1] * cy;
        b[2] = t[kPmlBa + 2] * b[2] - t[kPmlBb + 2] * cz;
    } else {
        b[0] = b[0] - dt * cx;
        b[1] = b[1] - dt * cy;
        b[2] = b[2] - dt * cz;
    }
}

// Update coefficients Ca, Cb of the three E edges of texel p (shaders/update_E.comp:46-61).
// PM = false compiles out the case of a material texel inside the absorber (the exponentials of
// its summed conductivity): only for lattices known to have none (LatticeArgs::pml_materials 0),
// where the kernel then needs fewer registers.
// Ca, Cb of a material texel (class id != 0) inside the absorber: its edge sigma plus the
// absorber's at the E stagger (shaders/update_E.comp:46-61).
__device__ __forceinline__ void absorber_material_coefficients(const LatticeArgs &a, int x, int y, int z, uint32_t id,
                                                               float ca[3], float cb[3]) {
    const float dt = a.dt;
    float sig[3] = {0.0f, 0.0f, 0.0f}, eps[3] = {1.0f, 1.0f, 1.0f};
    {
        const float *t = a.table + static_cast<size_t>(id) * kTableStride;
        sig[0] = t[kSig + 0]; sig[1] = t[kSig + 1]; sig[2] = t[kSig + 2];
        eps[0] = t[kEps + 0]; eps[1] = t[kEps + 1]; eps[2] = t[kEps + 2];
    }
    sig[0] += pml_sigma(a, x, 1, y, 0, z, 0);
    sig[1] += pml_sigma(a, x, 0, y, 1, z, 0);
    sig[2] += pml_sigma(a, x, 0, y, 0, z, 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float dtEps = dt / eps[c];
        const float xx = sig[c] * dtEps;
        ca[c] = det_expf(-xx);
        cb[c] = dtEps * phi_f(xx);
    }
}

template <bool PM = true>
__device__ __forceinline__ void edge_coefficients(const LatticeArgs &a, int x, int y, int z, uint32_t id,
                                                  float ca[3], float cb[3]) {
    const float dt = a.dt;
    if (in_pml(a, x, y, z) && id == 0u) {
        // vacuum edge inside the absorber: tabulated per depth class
        const float *t = pml_entry(a, x, y, z);
        ca[0] = t[kPmlEa + 0]; ca[1] = t[kPmlEa + 1]; ca[2] = t[kPmlEa + 2];
        cb[0] = t[kPmlEb + 0]; cb[1] = t[kPmlEb + 1]; cb[2] = t[kPmlEb + 2];
    } else if (PM && in_pml(a, x, y, z)) {
        absorber_material_coefficients(a, x, y, z, id, ca, cb);
    } else if (id != 0u) {
        const float *t = a.table + static_cast<size_t>(id) * kTableStride;
        ca[0] = t[kCa + 0]; ca[1] = t[kCa + 1]; ca[2] = t[kCa + 2];
        cb[0] = t[kCb + 0]; cb[1] = t[kCb + 1]; cb[2] = t[kCb + 2];
    } else {
        // vacuum (class 0 or no table): x = 0 -> exp(-0) = 1, phi(0) = 1, dt / 1 = dt
        ca[0] = ca[1] = ca[2] = 1.0f;
        cb[0] = cb[1] = cb[2] = dt;
    }
}


}  // namespace emfd

namespace emfd {

// One component of the Faraday update at texel q (shaders/update_B.comp:41-67), for kernels that
// rebuild a neighbour's B^{n+1/2} instead of exchanging it. t0 / t1 are the two curl terms of
// component c (term0 - term1); identical operation order to faraday().
__device__ __forceinline__ float faraday_component(const LatticeArgs &a, int c, int x, int y, int z, float t0, float t1,
                                                   uint32_t id, float b) {
    if (a.has_wire && id != 0u) {
        const float *w = a.table + static_cast<size_t>(id) * kTableStride + kWire + 2 * c;
        t0 *= w[0];
        t1 *= w[1];
    }
    const float cc = t0 - t1;
    const float dt = a.dt;
    if (in_pml(a, x, y, z)) {
        const float *t = pml_entry(a, x, y, z);
        return t[kPmlBa + c] * b - t[kPmlBb + c] * cc;
    }
    return b - dt * cc;
}

}  // namespace emfd
