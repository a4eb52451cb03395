// Device helpers of the field sampling and the port monitor, shared by the single-grid and
// batched launches. Compiled with -ffp-contract=off, like the Yee kernels.
#pragma once

#include "internal.h"

namespace emfd {

// One field value of component comp (0..2: E, 3..5: B) at lattice index i, any storage.
__device__ __forceinline__ float field_load(const FieldPtrs &f, int comp, size_t i) {
    const void *plane = comp < 3 ? static_cast<const void *>(f.e[comp]) : static_cast<const void *>(f.b[comp - 3]);
    if (f.storage == kStorageF32) return static_cast<const float *>(plane)[i];
    if (f.storage == kStorageF16) return static_cast<float>(static_cast<const _Float16 *>(plane)[i]) * f.scale[comp];
    if (f.storage == kStorageF64) return static_cast<float>(static_cast<const double *>(plane)[i]);
    return static_cast<float>(static_cast<const __bf16 *>(plane)[i]) * f.scale[comp];
}

// The value field_load returns after field_store(f, comp, i, v): v rounded through the storage
// format (binary16 / bfloat16 of v / scale, times scale; FP32 and FP64 hold v itself).
__device__ __forceinline__ float storage_value(const FieldPtrs &f, int comp, float v) {
    if (f.storage == kStorageF16) return static_cast<float>(static_cast<_Float16>(v / f.scale[comp])) * f.scale[comp];
    if (f.storage == kStorageBF16) return static_cast<float>(static_cast<__bf16>(v / f.scale[comp])) * f.scale[comp];
    return v;
}

__device__ __forceinline__ void field_store(const FieldPtrs &f, int comp, size_t i, float v) {
    void *plane = comp < 3 ? static_cast<void *>(f.e[comp]) : static_cast<void *>(f.b[comp - 3]);
    if (f.storage == kStorageF32) {
        static_cast<float *>(plane)[i] = v;
    } else if (f.storage == kStorageF16) {
        static_cast<_Float16 *>(plane)[i] = static_cast<_Float16>(v / f.scale[comp]);
    } else if (f.storage == kStorageF64) {
        static_cast<double *>(plane)[i] = static_cast<double>(v);
    } else {
        static_cast<__bf16 *>(plane)[i] = static_cast<__bf16>(v / f.scale[comp]);
    }
}

__device__ __forceinline__ float fetch_c(const LatticeArgs &a, const FieldPtrs &f, int comp, int x, int y, int z) {
    // texelFetch with indices clamped to the grid (shaders/field_sampling.glsl:14-15,
    // shaders/port_accumulate.comp:40-41)
    x = min(max(x, 0), a.nx - 1);
    y = min(max(y, 0), a.ny - 1);
    z = min(max(z, 0), a.nz - 1);
    return field_load(f, comp, (static_cast<size_t>(z) * a.ny + y) * a.nx + x);
}

// eAtNode / bAtNode (shaders/field_sampling.glsl:17-32) over any fetch(comp, x, y, z)
template <class Fetch>
__device__ __forceinline__ void e_at_node_with(const Fetch &fetch, int x, int y, int z, float o[3]) {
    o[0] = 0.5f * (fetch(0, x, y, z) + fetch(0, x - 1, y, z));
    o[1] = 0.5f * (fetch(1, x, y, z) + fetch(1, x, y - 1, z));
    o[2] = 0.5f * (fetch(2, x, y, z) + fetch(2, x, y, z - 1));
}
template <class Fetch>
__device__ __forceinline__ void b_at_node_with(const Fetch &fetch, int x, int y, int z, float o[3]) {
    const float bx = fetch(3, x, y, z) + fetch(3, x, y - 1, z) + fetch(3, x, y, z - 1) + fetch(3, x, y - 1, z - 1);
    const float by = fetch(4, x, y, z) + fetch(4, x - 1, y, z) + fetch(4, x, y, z - 1) + fetch(4, x - 1, y, z - 1);
    const float bz = fetch(5, x, y, z) + fetch(5, x - 1, y, z) + fetch(5, x, y - 1, z) + fetch(5, x - 1, y - 1, z);
    o[0] = 0.25f * bx;
    o[1] = 0.25f * by;
    o[2] = 0.25f * bz;
}
__device__ __forceinline__ void e_at_node(const LatticeArgs &a, const FieldPtrs &f, int x, int y, int z, float o[3]) {
    e_at_node_with([&](int c, int xx, int yy, int zz) { return fetch_c(a, f, c, xx, yy, zz); }, x, y, z, o);
}
__device__ __forceinline__ void b_at_node(const LatticeArgs &a, const FieldPtrs &f, int x, int y, int z, float o[3]) {
    b_at_node_with([&](int c, int xx, int yy, int zz) { return fetch_c(a, f, c, xx, yy, zz); }, x, y, z, o);
}

// ---------------------------------------------------------------------------------------------
// port_accumulate.comp (shaders/port_accumulate.comp:45-83)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float edge_current(const LatticeArgs &a, const FieldPtrs &f, int p[3], int ax) {
    const int u = (ax + 1) % 3, v = (ax + 2) % 3;
    int pu[3] = {p[0], p[1], p[2]}, pv[3] = {p[0], p[1], p[2]};
    pu[u] -= 1;
    pv[v] -= 1;
    const float b0v = fetch_c(a, f, 3 + v, p[0], p[1], p[2]), buv = fetch_c(a, f, 3 + v, pu[0], pu[1], pu[2]);
    const float b0u = fetch_c(a, f, 3 + u, p[0], p[1], p[2]), bvu = fetch_c(a, f, 3 + u, pv[0], pv[1], pv[2]);
    return (b0v - buv) - (b0u - bvu);
}

// A port's gap voltage and loop current (shaders/port_accumulate.comp:46-70): the thin gap's E
// and B loop, or the rod gap's four edges (V averaged, I summed); signed by the port's sign.
__device__ __forceinline__ void port_gap(const LatticeArgs &a, const FieldPtrs &f, const PortDev &pt, float *Vo,
                                         float *Io) {
    int p[3] = {pt.texel[0], pt.texel[1], pt.texel[2]};
    const int ax = pt.axis;
    float V = 0.0f, I = 0.0f;
    if (pt.edges == 4) {
        const int u = (ax + 1) % 3, v = (ax + 2) % 3;
        for (int k = 0; k < 4; ++k) {
            int q[3] = {p[0], p[1], p[2]};
            if (k == 1 || k == 3) q[u] += 1;
            if (k == 2 || k == 3) q[v] += 1;
            V += 0.25f * fetch_c(a, f, ax, q[0], q[1], q[2]);
            I += edge_current(a, f, q, ax);
        }
    } else {
        V = fetch_c(a, f, ax, p[0], p[1], p[2]);
        I = edge_current(a, f, p, ax);
    }
    const float s = static_cast<float>(pt.sign);
    *Vo = V * s;
    *Io = I * s;
}

// One port's gap sample and two-stage EMA update (shaders/port_accumulate.comp:46-83).
// ph: the port's 8 phasor floats; series_vi: where to record (V, I), or null.
__device__ __forceinline__ void port_sample(const LatticeArgs &a, const FieldPtrs &f, const PortDev &pt,
                                            const float *rot, float blend, float *ph, float *series_vi) {
    float V, I;
    port_gap(a, f, pt, &V, &I);
    if (series_vi) {
        series_vi[0] = V;
        series_vi[1] = I;
    }
    const float obs[4] = {2.0f * rot[0] * V, -2.0f * rot[1] * V, 2.0f * rot[2] * I, -2.0f * rot[3] * I};
    float *s1 = ph, *s2 = s1 + 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) s1[c] += blend * (obs[c] - s1[c]);
#pragma unroll
    for (int c = 0; c < 4; ++c) s2[c] += blend * (s1[c] - s2[c]);
}

}  // namespace emfd
