// Per-quad bodies of the vectorised Yee passes: one thread owns four consecutive x cells so
// every field stream moves as float4 (16 B per lane). Requires nx % 4 == 0 and x0 % 4 == 0.
// The x+1 (B pass) and x-1 (E pass) neighbours of the edge cells are extra dword loads that hit
// the same cache lines. Same per-texel arithmetic as faraday() and edge_coefficients(), so every
// kernel built on these bodies (single grid or batched) gives bit-identical fields.
#pragma once

#include "coef_rows.h"

namespace emfd {

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float a, float b, float c, float d) {
    *reinterpret_cast<float4 *>(p) = make_float4(a, b, c, d);
}
// Streaming forms (non-temporal hint): for the plane a pass reads or writes exactly once.
__device__ __forceinline__ float4 ld4s(const float *p) {
    const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st4s(float *p, float a, float b, float c, float d) {
    const f32x4 v = {a, b, c, d};
    __builtin_nontemporal_store(v, reinterpret_cast<f32x4 *>(p));
}
__device__ __forceinline__ void unpack(const float4 v, float (&o)[4]) {
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// Copies the four cells' values of components comp0..comp0+2 into the far-field side columns when
// the quad holds one of them.
// (per-cell side_index() measured 0.8% faster here than side_write_cells(), which pays off for the
// 8-cell runs of the 16-bit kernels)
__device__ __forceinline__ void side_write4(const LatticeArgs &a, int comp0, int x0, int y, int z,
                                            const float (&v)[3][4]) {
    const SideCols &s = a.side;
    if (!s.buf || y < s.y0 || y >= s.y0 + s.ny || z < s.z0 || z >= s.z0 + s.nz) return;
    if (x0 + 3 < s.x[0] - 1 || (x0 > s.x[0] && x0 + 3 < s.x[1] - 1) || x0 > s.x[1]) return;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int face = 0; face < 2; ++face) {
            const long k = side_index(s, face, comp0, x0 + j, y, z);
            if (k < 0) continue;
            const long plane = 2L * s.nz * s.nyp;  // component stride
#pragma unroll
            for (int c = 0; c < 3; ++c) s.buf[k + c * plane] = v[c][j];
        }
}

template <bool STREAM = false>
__device__ __forceinline__ void load_ids4(const LatticeArgs &a, size_t p, uint32_t (&id)[4]) {
    if (!a.mat_id) {
        id[0] = id[1] = id[2] = id[3] = 0u;
    } else if (a.id_bytes == 2) {
        uint2 v;
        if (STREAM) {
            const u32x2 w = __builtin_nontemporal_load(
                reinterpret_cast<const u32x2 *>(static_cast<const uint16_t *>(a.mat_id) + p));
            v = make_uint2(w.x, w.y);
        } else {
            v = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(a.mat_id) + p);
        }
        id[0] = v.x & 0xffffu; id[1] = v.x >> 16; id[2] = v.y & 0xffffu; id[3] = v.y >> 16;
    } else {
        const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(a.mat_id) + p);
        id[0] = v.x; id[1] = v.y; id[2] = v.z; id[3] = v.w;
    }
}

// update_B for the four cells x0..x0+3 of row (y, z) (shaders/update_B.comp:31-70). ST: 0 plain
// memory operations, 1 streaming stores, 2 streaming stores and streaming loads of the B plane
// (read once per pass), 3 as 2 plus streaming loads of the class ids; the arithmetic is the same
// for all.
template <int ST = 0>
__device__ __forceinline__ void b_quad(const LatticeArgs &a, const FieldPtrs &f, int x0, int y, int z) {
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x0;
    const bool up_x = x0 + 4 < a.nx, up_y = y + 1 < a.ny, up_z = z + 1 < a.nz;
    float e0[3][4], eX[3], eY[3][4], eZ[3][4], b[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *e = f.e[c];
        unpack(ld4(e + p), e0[c]);
        eX[c] = up_x ? e[p + 4] : 0.0f;
        if (up_y) unpack(ld4(e + p + sy), eY[c]); else eY[c][0] = eY[c][1] = eY[c][2] = eY[c][3] = 0.0f;
        if (up_z) unpack(ld4(e + p + sz), eZ[c]); else eZ[c][0] = eZ[c][1] = eZ[c][2] = eZ[c][3] = 0.0f;
        unpack(ST >= 2 ? ld4s(f.b[c] + p) : ld4(f.b[c] + p), b[c]);
    }
    // the Faraday update reads a class only for its thin-wire factors: no class stream at all
    // without wires, and with a wire mask only the quads that hold a wire texel load theirs
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    if (a.has_wire) {
        if (a.wire_mask) {
            const uint32_t w = a.wire_mask[(static_cast<size_t>(z) * a.ny + y) * a.mask_wpr + (x0 >> 5)];
            if ((w >> (x0 & 31)) & 0xfu) load_ids4<(ST >= 3)>(a, p, id);
        } else {
            load_ids4<(ST >= 3)>(a, p, id);
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float c0[3] = {e0[0][j], e0[1][j], e0[2][j]};
        const float cx[3] = {j < 3 ? e0[0][j + 1] : eX[0], j < 3 ? e0[1][j + 1] : eX[1], j < 3 ? e0[2][j + 1] : eX[2]};
        const float cy[3] = {eY[0][j], eY[1][j], eY[2][j]};
        const float cz[3] = {eZ[0][j], eZ[1][j], eZ[2][j]};
        float bb[3] = {b[0][j], b[1][j], b[2][j]};
        faraday(a, x0 + j, y, z, c0, cx, cy, cz, id[j], bb);
        b[0][j] = bb[0];
        b[1][j] = bb[1];
        b[2][j] = bb[2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (ST >= 1) st4s(f.b[c] + p, b[c][0], b[c][1], b[c][2], b[c][3]);
        else st4(f.b[c] + p, b[c][0], b[c][1], b[c][2], b[c][3]);
    }
    side_write4(a, 3, x0, y, z, b);
}

// update_E for the four cells x0..x0+3 of row (y, z) (shaders/update_E.comp:31-66); ST as b_quad
// with the E plane as the once-read plane; PM as edge_coefficients.
template <int ST = 0, bool PM = true>
__device__ __forceinline__ void e_quad(const LatticeArgs &a, const FieldPtrs &f, int x0, int y, int z) {
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x0;
    const bool dn_x = x0 > 0, dn_y = y > 0, dn_z = z > 0;
    float b0[3][4], bX[3], bY[3][4], bZ[3][4], e[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *bp = f.b[c];
        unpack(ld4(bp + p), b0[c]);
        bX[c] = dn_x ? bp[p - 1] : 0.0f;
        if (dn_y) unpack(ld4(bp + p - sy), bY[c]); else bY[c][0] = bY[c][1] = bY[c][2] = bY[c][3] = 0.0f;
        if (dn_z) unpack(ld4(bp + p - sz), bZ[c]); else bZ[c][0] = bZ[c][1] = bZ[c][2] = bZ[c][3] = 0.0f;
        unpack(ST >= 2 ? ld4s(f.e[c] + p) : ld4(f.e[c] + p), e[c]);
    }
    // class records only where the quad holds a texel of class != 0 (the class mask says which)
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    if (a.class_mask) {
        const uint32_t w = a.class_mask[(static_cast<size_t>(z) * a.ny + y) * a.mask_wpr + (x0 >> 5)];
        if ((w >> (x0 & 31)) & 0xfu) load_ids4<(ST >= 3)>(a, p, id);
    } else {
        load_ids4<(ST >= 3)>(a, p, id);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float bxm[3] = {j > 0 ? b0[0][j - 1] : bX[0], j > 0 ? b0[1][j - 1] : bX[1], j > 0 ? b0[2][j - 1] : bX[2]};
        const float curl[3] = {(b0[2][j] - bY[2][j]) - (b0[1][j] - bZ[1][j]), (b0[0][j] - bZ[0][j]) - (b0[2][j] - bxm[2]),
                               (b0[1][j] - bxm[1]) - (b0[0][j] - bY[0][j])};
        float ca[3], cb[3];
        edge_coefficients<PM>(a, x0 + j, y, z, id[j], ca, cb);
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c][j] = ca[c] * e[c][j] + cb[c] * curl[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (ST >= 1) st4s(f.e[c] + p, e[c][0], e[c][1], e[c][2], e[c][3]);
        else st4(f.e[c] + p, e[c][0], e[c][1], e[c][2], e[c][3]);
    }
    side_write4(a, 0, x0, y, z, e);
}

// b_quad() for a kernel whose wavefront updates one row (64 lanes x 4 cells) with its absorber
// coefficients staged in LDS (coef_rows.h): `rows` is the wavefront's LDS region, `lane` its lane.
// Lanes beyond the lattice (!active) load at cell (0, y0, z) and take part in the staging only.
// Quads without a thin-wire texel take one branch-free path (faraday_coef); the others faraday().
template <int ST = 0>
__device__ __forceinline__ void b_quad_rows(const LatticeArgs &a, const FieldPtrs &f, int x0, int y, int z, bool active,
                                            float *rows, int lane) {
    const CoefLoad coef0 = coef_load<1>(a, lane, y, z, kPmlBa);
    const int xa = active ? x0 : 0, ya = y < a.ny ? y : 0;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + ya * sy + xa;
    const bool up_x = xa + 4 < a.nx, up_y = ya + 1 < a.ny, up_z = z + 1 < a.nz;
    float e0[3][4], eX[3], eY[3][4], eZ[3][4], b[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *e = f.e[c];
        unpack(ld4(e + p), e0[c]);
        eX[c] = up_x ? e[p + 4] : 0.0f;
        if (up_y) unpack(ld4(e + p + sy), eY[c]); else eY[c][0] = eY[c][1] = eY[c][2] = eY[c][3] = 0.0f;
        if (up_z) unpack(ld4(e + p + sz), eZ[c]); else eZ[c][0] = eZ[c][1] = eZ[c][2] = eZ[c][3] = 0.0f;
        unpack(ST >= 2 ? ld4s(f.b[c] + p) : ld4(f.b[c] + p), b[c]);
    }
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    if (a.has_wire) {
        if (a.wire_mask) {
            const uint32_t w = a.wire_mask[(static_cast<size_t>(z) * a.ny + ya) * a.mask_wpr + (xa >> 5)];
            if ((w >> (xa & 31)) & 0xfu) load_ids4<(ST >= 3)>(a, p, id);
        } else {
            load_ids4<(ST >= 3)>(a, p, id);
        }
    }
    coef_store(a, rows, lane, coef0);
    coef_finish<1>(a, rows, lane, y, z, kPmlBa);
    if (!active) return;
    if ((id[0] | id[1] | id[2] | id[3]) == 0u) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float c0[3] = {e0[0][j], e0[1][j], e0[2][j]};
            const float cx[3] = {j < 3 ? e0[0][j + 1] : eX[0], j < 3 ? e0[1][j + 1] : eX[1], j < 3 ? e0[2][j + 1] : eX[2]};
            const float cy[3] = {eY[0][j], eY[1][j], eY[2][j]};
            const float cz[3] = {eZ[0][j], eZ[1][j], eZ[2][j]};
            float bb[3] = {b[0][j], b[1][j], b[2][j]}, ka[3], kb[3];
            read_coef(rows + pml_class(x0 + j, a.nx, a.pml) * kCoefStride, ka, kb);
            faraday_coef(a, c0, cx, cy, cz, 0u, ka, kb, bb);
            b[0][j] = bb[0];
            b[1][j] = bb[1];
            b[2][j] = bb[2];
        }
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float c0[3] = {e0[0][j], e0[1][j], e0[2][j]};
            const float cx[3] = {j < 3 ? e0[0][j + 1] : eX[0], j < 3 ? e0[1][j + 1] : eX[1], j < 3 ? e0[2][j + 1] : eX[2]};
            const float cy[3] = {eY[0][j], eY[1][j], eY[2][j]};
            const float cz[3] = {eZ[0][j], eZ[1][j], eZ[2][j]};
            float bb[3] = {b[0][j], b[1][j], b[2][j]};
            faraday(a, x0 + j, y, z, c0, cx, cy, cz, id[j], bb);
            b[0][j] = bb[0];
            b[1][j] = bb[1];
            b[2][j] = bb[2];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (ST >= 1) st4s(f.b[c] + p, b[c][0], b[c][1], b[c][2], b[c][3]);
        else st4(f.b[c] + p, b[c][0], b[c][1], b[c][2], b[c][3]);
    }
    side_write4(a, 3, x0, y, z, b);
}

// e_quad() with staged absorber coefficients, as b_quad_rows(): vacuum edges read theirs from
// the wavefront's LDS rows; material edges as edge_coefficients().
template <int ST = 0, bool PM = true>
__device__ __forceinline__ void e_quad_rows(const LatticeArgs &a, const FieldPtrs &f, int x0, int y, int z, bool active,
                                            float *rows, int lane) {
    const CoefLoad coef0 = coef_load<1>(a, lane, y, z, kPmlEa);
    const int xa = active ? x0 : 0, ya = y < a.ny ? y : 0;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + ya * sy + xa;
    const bool dn_x = xa > 0, dn_y = ya > 0, dn_z = z > 0;
    float b0[3][4], bX[3], bY[3][4], bZ[3][4], e[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *bp = f.b[c];
        unpack(ld4(bp + p), b0[c]);
        bX[c] = dn_x ? bp[p - 1] : 0.0f;
        if (dn_y) unpack(ld4(bp + p - sy), bY[c]); else bY[c][0] = bY[c][1] = bY[c][2] = bY[c][3] = 0.0f;
        if (dn_z) unpack(ld4(bp + p - sz), bZ[c]); else bZ[c][0] = bZ[c][1] = bZ[c][2] = bZ[c][3] = 0.0f;
        unpack(ST >= 2 ? ld4s(f.e[c] + p) : ld4(f.e[c] + p), e[c]);
    }
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    if (a.class_mask) {
        const uint32_t w = a.class_mask[(static_cast<size_t>(z) * a.ny + ya) * a.mask_wpr + (xa >> 5)];
        if ((w >> (xa & 31)) & 0xfu) load_ids4<(ST >= 3)>(a, p, id);
    } else {
        load_ids4<(ST >= 3)>(a, p, id);
    }
    coef_store(a, rows, lane, coef0);
    coef_finish<1>(a, rows, lane, y, z, kPmlEa);
    if (!active) return;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float bxm[3] = {j > 0 ? b0[0][j - 1] : bX[0], j > 0 ? b0[1][j - 1] : bX[1], j > 0 ? b0[2][j - 1] : bX[2]};
        const float curl[3] = {(b0[2][j] - bY[2][j]) - (b0[1][j] - bZ[1][j]), (b0[0][j] - bZ[0][j]) - (b0[2][j] - bxm[2]),
                               (b0[1][j] - bxm[1]) - (b0[0][j] - bY[0][j])};
        float ca[3], cb[3];
        // every texel reads its staged entry; a material texel then replaces it (read first and
        // unconditionally, the vacuum entry does not wait behind the material texels' class loads)
        read_coef(rows + pml_class(x0 + j, a.nx, a.pml) * kCoefStride, ca, cb);
        if (id[j] != 0u) {
            if (PM && in_pml(a, x0 + j, y, z)) {
                absorber_material_coefficients(a, x0 + j, y, z, id[j], ca, cb);
            } else {
                const float *t = a.table + static_cast<size_t>(id[j]) * kTableStride;
                ca[0] = t[kCa + 0]; ca[1] = t[kCa + 1]; ca[2] = t[kCa + 2];
                cb[0] = t[kCb + 0]; cb[1] = t[kCb + 1]; cb[2] = t[kCb + 2];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c][j] = ca[c] * e[c][j] + cb[c] * curl[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (ST >= 1) st4s(f.e[c] + p, e[c][0], e[c][1], e[c][2], e[c][3]);
        else st4(f.e[c] + p, e[c][0], e[c][1], e[c][2], e[c][3]);
    }
    side_write4(a, 0, x0, y, z, e);
}

// The four cells' absorber table entries, the B half (OFF = kPmlBa: exp(-s dt), dt phi(s dt)) or
// the E half (OFF = kPmlEa: Ca, Cb of a vacuum edge), loaded unconditionally with the field
// loads. Entry 0 (no absorber on any axis) holds the interior coefficients (1, dt) (the table
// builder writes it so), which give the interior update bit for bit; behind a per-cell branch
// each cell of an absorber quad waited for its own table load in turn.
template <int OFF>
__device__ __forceinline__ void quad_coefs(const LatticeArgs &a, int x0, int y, int z, float (&lo)[4][3],
                                           float (&hi)[4][3]) {
    const int K = a.pml_k, L = a.pml;
    const int cyz = K * (pml_class(y, a.ny, L) + K * pml_class(z, a.nz, L));
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float *t = a.pml_tab + static_cast<size_t>(pml_class(x0 + j, a.nx, L) + cyz) * kPmlStride + OFF;
        if (OFF == kPmlBa) {
            const float4 u = *reinterpret_cast<const float4 *>(t);
            const float2 w = *reinterpret_cast<const float2 *>(t + 4);
            lo[j][0] = u.x; lo[j][1] = u.y; lo[j][2] = u.z;
            hi[j][0] = u.w; hi[j][1] = w.x; hi[j][2] = w.y;
        } else {
            const float2 u = *reinterpret_cast<const float2 *>(t);
            const float2 v = *reinterpret_cast<const float2 *>(t + 2);
            const float2 w = *reinterpret_cast<const float2 *>(t + 4);
            lo[j][0] = u.x; lo[j][1] = u.y; lo[j][2] = v.x;
            hi[j][0] = v.y; hi[j][1] = w.x; hi[j][2] = w.y;
        }
    }
}

// b_quad() with the absorber coefficients loaded up front (quad_coefs): quads without a thin-wire
// texel take one branch-free path, the others faraday(). For kernels whose wavefronts do not
// cover whole rows (the batched stepper).
template <int ST = 0>
__device__ __forceinline__ void b_quad_pre(const LatticeArgs &a, const FieldPtrs &f, int x0, int y, int z) {
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x0;
    const bool up_x = x0 + 4 < a.nx, up_y = y + 1 < a.ny, up_z = z + 1 < a.nz;
    float e0[3][4], eX[3], eY[3][4], eZ[3][4], b[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *e = f.e[c];
        unpack(ld4(e + p), e0[c]);
        eX[c] = up_x ? e[p + 4] : 0.0f;
        if (up_y) unpack(ld4(e + p + sy), eY[c]); else eY[c][0] = eY[c][1] = eY[c][2] = eY[c][3] = 0.0f;
        if (up_z) unpack(ld4(e + p + sz), eZ[c]); else eZ[c][0] = eZ[c][1] = eZ[c][2] = eZ[c][3] = 0.0f;
        unpack(ST >= 2 ? ld4s(f.b[c] + p) : ld4(f.b[c] + p), b[c]);
    }
    float ka[4][3], kb[4][3];
    quad_coefs<kPmlBa>(a, x0, y, z, ka, kb);
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    if (a.has_wire) {
        if (a.wire_mask) {
            const uint32_t w = a.wire_mask[(static_cast<size_t>(z) * a.ny + y) * a.mask_wpr + (x0 >> 5)];
            if ((w >> (x0 & 31)) & 0xfu) load_ids4<(ST >= 3)>(a, p, id);
        } else {
            load_ids4<(ST >= 3)>(a, p, id);
        }
    }
    const bool wired = (id[0] | id[1] | id[2] | id[3]) != 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float c0[3] = {e0[0][j], e0[1][j], e0[2][j]};
        const float cx[3] = {j < 3 ? e0[0][j + 1] : eX[0], j < 3 ? e0[1][j + 1] : eX[1], j < 3 ? e0[2][j + 1] : eX[2]};
        const float cy[3] = {eY[0][j], eY[1][j], eY[2][j]};
        const float cz[3] = {eZ[0][j], eZ[1][j], eZ[2][j]};
        float bb[3] = {b[0][j], b[1][j], b[2][j]};
        if (!wired) faraday_coef(a, c0, cx, cy, cz, 0u, ka[j], kb[j], bb);
        else faraday(a, x0 + j, y, z, c0, cx, cy, cz, id[j], bb);
        b[0][j] = bb[0];
        b[1][j] = bb[1];
        b[2][j] = bb[2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (ST >= 1) st4s(f.b[c] + p, b[c][0], b[c][1], b[c][2], b[c][3]);
        else st4(f.b[c] + p, b[c][0], b[c][1], b[c][2], b[c][3]);
    }
    side_write4(a, 3, x0, y, z, b);
}

// e_quad() with the vacuum edges' absorber (or interior) coefficients loaded up front; material
// edges as edge_coefficients().
template <int ST = 0, bool PM = true>
__device__ __forceinline__ void e_quad_pre(const LatticeArgs &a, const FieldPtrs &f, int x0, int y, int z) {
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x0;
    const bool dn_x = x0 > 0, dn_y = y > 0, dn_z = z > 0;
    float b0[3][4], bX[3], bY[3][4], bZ[3][4], e[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *bp = f.b[c];
        unpack(ld4(bp + p), b0[c]);
        bX[c] = dn_x ? bp[p - 1] : 0.0f;
        if (dn_y) unpack(ld4(bp + p - sy), bY[c]); else bY[c][0] = bY[c][1] = bY[c][2] = bY[c][3] = 0.0f;
        if (dn_z) unpack(ld4(bp + p - sz), bZ[c]); else bZ[c][0] = bZ[c][1] = bZ[c][2] = bZ[c][3] = 0.0f;
        unpack(ST >= 2 ? ld4s(f.e[c] + p) : ld4(f.e[c] + p), e[c]);
    }
    float ka[4][3], kb[4][3];
    quad_coefs<kPmlEa>(a, x0, y, z, ka, kb);
    uint32_t id[4] = {0u, 0u, 0u, 0u};
    if (a.class_mask) {
        const uint32_t w = a.class_mask[(static_cast<size_t>(z) * a.ny + y) * a.mask_wpr + (x0 >> 5)];
        if ((w >> (x0 & 31)) & 0xfu) load_ids4<(ST >= 3)>(a, p, id);
    } else {
        load_ids4<(ST >= 3)>(a, p, id);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float bxm[3] = {j > 0 ? b0[0][j - 1] : bX[0], j > 0 ? b0[1][j - 1] : bX[1], j > 0 ? b0[2][j - 1] : bX[2]};
        const float curl[3] = {(b0[2][j] - bY[2][j]) - (b0[1][j] - bZ[1][j]), (b0[0][j] - bZ[0][j]) - (b0[2][j] - bxm[2]),
                               (b0[1][j] - bxm[1]) - (b0[0][j] - bY[0][j])};
        float ca[3] = {ka[j][0], ka[j][1], ka[j][2]}, cb[3] = {kb[j][0], kb[j][1], kb[j][2]};
        if (id[j] != 0u) {
            if (PM && in_pml(a, x0 + j, y, z)) {
                absorber_material_coefficients(a, x0 + j, y, z, id[j], ca, cb);
            } else {
                const float *t = a.table + static_cast<size_t>(id[j]) * kTableStride;
                ca[0] = t[kCa + 0]; ca[1] = t[kCa + 1]; ca[2] = t[kCa + 2];
                cb[0] = t[kCb + 0]; cb[1] = t[kCb + 1]; cb[2] = t[kCb + 2];
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c][j] = ca[c] * e[c][j] + cb[c] * curl[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (ST >= 1) st4s(f.e[c] + p, e[c][0], e[c][1], e[c][2], e[c][3]);
        else st4(f.e[c] + p, e[c][0], e[c][1], e[c][2], e[c][3]);
    }
    side_write4(a, 0, x0, y, z, e);
}

}  // namespace emfd
