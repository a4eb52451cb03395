// Absorber coefficients staged in LDS per wavefront, for the Yee passes whose wavefronts each
// update whole x runs of one or two rows (R): the shared device pieces of the FP32 and 16-bit
// two-pass kernels.
#pragma once

#include "yee_device.h"

namespace emfd {

// Absorber coefficients of the rows a wavefront updates, staged in LDS by that wavefront: entry
// (row r, x depth class cx) holds the six floats at offset `off` of the depth-class entry
// (cx, class(y_r), class(z)) (B: exp(-s dt) x3, dt phi(s dt) x3; E: Ca x3, Cb x3 of a vacuum
// edge). The entry of a texel outside the absorber holds (1, 1, 1, dt, dt, dt): 1 * b - dt * c
// equals b - dt * c exactly, and (1, dt) are the vacuum edge's Ca, Cb, so every texel of a row
// takes one branch-free path. A row spans the x layers at both ends, so without this every
// wavefront ran the absorber branch (and waited on its table load) for each of its cells.
constexpr int kCoefStride = 8;  // floats per LDS entry (6 used)

// entries per wavefront region: its rows' entries, at least one per lane
__host__ __device__ __forceinline__ int coef_entries(const LatticeArgs &a, int rows) {
    return rows * a.pml_k > 64 ? rows * a.pml_k : 64;
}

// One LDS entry of the staged rows: entry i = r * K + cx. The table load is unconditional (at a
// valid entry) and the interior value is selected after it, so the wavefront issues these loads
// together with its field loads and waits once.
struct CoefLoad {
    float v[6];
    bool interior;
};

template <int R>
__device__ __forceinline__ CoefLoad coef_load(const LatticeArgs &a, int i, int y_first, int z, int off) {
    const int K = a.pml_k, L = a.pml;
    CoefLoad c;
    const int ii = i < R * K ? i : 0;
    const int r = (R == 2 && ii >= K) ? 1 : 0;
    const int cx = ii - r * K;
    const int y = y_first + r;
    // a row or plane outside the lattice (idle work, nothing stored) takes class 0: a depth class
    // of its own would index past the table
    const int cy = y >= 0 && y < a.ny ? pml_class(y, a.ny, L) : 0, cz = z >= 0 && z < a.nz ? pml_class(z, a.nz, L) : 0;
    c.interior = (cx | cy | cz) == 0;
    const float *t = a.pml_tab + static_cast<size_t>(cx + K * (cy + K * cz)) * kPmlStride + off;
#pragma unroll
    for (int k = 0; k < 6; ++k) c.v[k] = t[k];
    return c;
}

// (unconditional: a wavefront's region holds at least 64 entries, so lanes past the last entry
// write spare ones; a conditional store let the compiler sink the table load behind the field
// loads)
__device__ __forceinline__ void coef_store(const LatticeArgs &a, float *lds, int i, const CoefLoad &c) {
    float v[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) v[k] = c.interior ? (k < 3 ? 1.0f : a.dt) : c.v[k];
    float *d = lds + i * kCoefStride;
    *reinterpret_cast<float4 *>(d) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float2 *>(d + 4) = make_float2(v[4], v[5]);
}

// Entries beyond the first 64 (absorbers thicker than 31 cells), then the wavefront-level
// ordering of the entry writes before any entry read: LDS operations of one wavefront complete
// in order, and the fences keep the compiler from moving the reads above the writes.
template <int R>
__device__ __forceinline__ void coef_finish(const LatticeArgs &a, float *lds, int lane, int y_first, int z, int off) {
    for (int i = lane + 64; i < R * a.pml_k; i += 64) coef_store(a, lds, i, coef_load<R>(a, i, y_first, z, off));
    (void)y_first;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void read_coef(const float *t, float (&lo)[3], float (&hi)[3]) {
    const float4 u = *reinterpret_cast<const float4 *>(t);
    const float2 w = *reinterpret_cast<const float2 *>(t + 4);
    lo[0] = u.x; lo[1] = u.y; lo[2] = u.z;
    hi[0] = u.w; hi[1] = w.x; hi[2] = w.y;
}

// faraday() with the texel's absorber coefficients given (ba, bb): the same operations as its
// absorber branch, which for (1, dt) are those of its interior branch
__device__ __forceinline__ void faraday_coef(const LatticeArgs &a, const float e0[3], const float exn[3],
                                             const float eyn[3], const float ezn[3], uint32_t id, const float ba[3],
                                             const float bb[3], float b[3]) {
    float t0x = eyn[2] - e0[2], t0y = ezn[0] - e0[0], t0z = exn[1] - e0[1];
    float t1x = ezn[1] - e0[1], t1y = exn[2] - e0[2], t1z = eyn[0] - e0[0];
    if (a.has_wire && id != 0u) {
        const float *w = a.table + static_cast<size_t>(id) * kTableStride + kWire;
        t0x *= w[0];
        t1x *= w[1];
        t0y *= w[2];
        t1y *= w[3];
        t0z *= w[4];
        t1z *= w[5];
    }
    const float cx = t0x - t1x, cy = t0y - t1y, cz = t0z - t1z;
    b[0] = ba[0] * b[0] - bb[0] * cx;
    b[1] = ba[1] * b[1] - bb[1] * cy;
    b[2] = ba[2] * b[2] - bb[2] * cz;
}

}  // namespace emfd
