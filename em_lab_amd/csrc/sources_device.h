// Device body of the source pass, shared by the single-grid and batched launches.
#pragma once

#include "port_device.h"
#include "yee_device.h"

namespace emfd {

// ---------------------------------------------------------------------------------------------
// apply_sources.comp (shaders/apply_sources.comp:37-79): one thread per affected texel.
// packed: 16 floats per source (center, halfExtent, polAmp, softness; vec4 each).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ float footprint(const float *s, float x, float y, float z) {
    const float dx = fmaxf(fabsf(x - s[0]) - s[4], 0.0f) / s[12];
    const float dy = fmaxf(fabsf(y - s[1]) - s[5], 0.0f) / s[13];
    const float dz = fmaxf(fabsf(z - s[2]) - s[6], 0.0f) / s[14];
    return det_expf(-(dx * dx + dy * dy + dz * dz));
}

// Source injection at source cell `id` of the per-texel list (cells: {x, y, z, first} with a
// sentinel; lists: source indices; signals: one value per source for this step).
__device__ __forceinline__ void source_cell(const LatticeArgs &a, const FieldPtrs &f, const int *cells,
                                            const int *lists, const float *packed, const float *signals, int id) {
    const int x = cells[4 * id + 0], y = cells[4 * id + 1], z = cells[4 * id + 2];
    const int first = cells[4 * id + 3], end = cells[4 * (id + 1) + 3];
    const float cx = static_cast<float>(x), cy = static_cast<float>(y), cz = static_cast<float>(z);
    float J[3] = {0.0f, 0.0f, 0.0f};
    for (int k = first; k < end; ++k) {
        const int i = lists[k];
        const float signal = signals[i];
        if (signal == 0.0f) continue;
        const float *s = packed + 16 * i;
        const float w[3] = {footprint(s, cx + 0.5f, cy + 0.0f, cz + 0.0f), footprint(s, cx + 0.0f, cy + 0.5f, cz + 0.0f),
                            footprint(s, cx + 0.0f, cy + 0.0f, cz + 0.5f)};
        const float as = s[11] * signal;
#pragma unroll
        for (int c = 0; c < 3; ++c) J[c] += s[8 + c] * as * w[c];
    }
    if (J[0] == 0.0f && J[1] == 0.0f && J[2] == 0.0f) return;
    const size_t p = (static_cast<size_t>(z) * a.ny + y) * a.nx + x;
    const uint32_t cid = a.mat_id ? load_id(a, p) : 0u;
    float ca[3], cb[3];
    edge_coefficients(a, x, y, z, cid, ca, cb);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float v = field_load(f, c, p) - cb[c] * J[c];
        field_store(f, c, p, v);
        if (a.side.buf)
            for (int face = 0; face < 2; ++face) {
                const long k = side_index(a.side, face, c, x, y, z);
                if (k >= 0) a.side.buf[k] = storage_value(f, c, v);  // what the planes now hold
            }
    }
}

// Per source texel, the parts of source_cell() that are fixed for a scene: the texel, its E-edge
// Cb (edge_coefficients) and the footprint weights of its first source; later sources of the
// texel (rare) in the pair list. Built by source_record() whenever sources or materials change;
// source_apply() then needs the record, the step's signals and the texel's fields, and computes
// what source_cell() computes, in the same order.
struct SourceRec {
    int x, y, z, first, end;  // texel; its sources: lists[first .. end-1] (pairs at the same indices)
    int i0;                   // lists[first]
    float cb[3];              // Cb of the texel's three E edges
    float w0[3];              // footprint weights of source i0 at the three E staggers
    int pad[4];
};
struct SourcePair {
    int i;       // source index
    float w[3];  // its footprint weights at this texel
};

__device__ __forceinline__ void source_weights(const float *s, float cx, float cy, float cz, float (&w)[3]) {
    w[0] = footprint(s, cx + 0.5f, cy + 0.0f, cz + 0.0f);
    w[1] = footprint(s, cx + 0.0f, cy + 0.5f, cz + 0.0f);
    w[2] = footprint(s, cx + 0.0f, cy + 0.0f, cz + 0.5f);
}

__device__ __forceinline__ void source_record(const LatticeArgs &a, const int *cells, const int *lists,
                                              const float *packed, int id, SourceRec *recs, SourcePair *pairs) {
    SourceRec r{};
    r.x = cells[4 * id + 0];
    r.y = cells[4 * id + 1];
    r.z = cells[4 * id + 2];
    r.first = cells[4 * id + 3];
    r.end = cells[4 * (id + 1) + 3];
    const float cx = static_cast<float>(r.x), cy = static_cast<float>(r.y), cz = static_cast<float>(r.z);
    for (int k = r.first; k < r.end; ++k) {
        SourcePair pr;
        pr.i = lists[k];
        source_weights(packed + 16 * pr.i, cx, cy, cz, pr.w);
        pairs[k] = pr;
        if (k == r.first) {
            r.i0 = pr.i;
            r.w0[0] = pr.w[0];
            r.w0[1] = pr.w[1];
            r.w0[2] = pr.w[2];
        }
    }
    const size_t p = (static_cast<size_t>(r.z) * a.ny + r.y) * a.nx + r.x;
    const uint32_t cid = a.mat_id ? load_id(a, p) : 0u;
    float ca[3];
    edge_coefficients(a, r.x, r.y, r.z, cid, ca, r.cb);
    recs[id] = r;
}

// source_cell() from a record: one dependent load (the record) before the signal and field loads
__device__ __forceinline__ void source_apply(const LatticeArgs &a, const FieldPtrs &f, const SourceRec *recs,
                                             const SourcePair *pairs, const float *packed, const float *signals,
                                             int id) {
    const SourceRec r = recs[id];
    if (r.first >= r.end) return;
    const size_t p = (static_cast<size_t>(r.z) * a.ny + r.y) * a.nx + r.x;
    const float e[3] = {field_load(f, 0, p), field_load(f, 1, p), field_load(f, 2, p)};
    float J[3] = {0.0f, 0.0f, 0.0f};
    {
        const float signal = signals[r.i0];
        if (signal != 0.0f) {
            const float *s = packed + 16 * r.i0;
            const float as = s[11] * signal;
#pragma unroll
            for (int c = 0; c < 3; ++c) J[c] += s[8 + c] * as * r.w0[c];
        }
    }
    for (int k = r.first + 1; k < r.end; ++k) {
        const SourcePair pr = pairs[k];
        const float signal = signals[pr.i];
        if (signal == 0.0f) continue;
        const float *s = packed + 16 * pr.i;
        const float as = s[11] * signal;
#pragma unroll
        for (int c = 0; c < 3; ++c) J[c] += s[8 + c] * as * pr.w[c];
    }
    if (J[0] == 0.0f && J[1] == 0.0f && J[2] == 0.0f) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float v = e[c] - r.cb[c] * J[c];
        field_store(f, c, p, v);
        if (a.side.buf)
            for (int face = 0; face < 2; ++face) {
                const long k = side_index(a.side, face, c, r.x, r.y, r.z);
                if (k >= 0) a.side.buf[k] = storage_value(f, c, v);  // what the planes now hold
            }
    }
}

}  // namespace emfd
