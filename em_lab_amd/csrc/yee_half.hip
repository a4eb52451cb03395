// Two-pass leapfrog on 16-bit field storage: E and B planes hold binary16 or bfloat16 values of
// field / scale (a power-of-two scale per component), every update runs in FP32 with the same
// per-texel arithmetic as faraday() / edge_coefficients(), and the result is rounded to nearest
// even on the way back. Each thread owns eight consecutive x cells so every field stream moves
// as 16-byte loads and stores (8 values). Requires nx % 8 == 0. Field traffic per cell-update is
// 24 B (half of FP32), plus the x+-1 neighbours of the edge cells, which hit the same lines.
#include "yee_device.h"

namespace emfd {
namespace {

template <typename T>
__device__ __forceinline__ void load8(const void *plane, size_t p, float s, float (&o)[8]) {
    const uint4 raw = *reinterpret_cast<const uint4 *>(static_cast<const T *>(plane) + p);
    const T *h = reinterpret_cast<const T *>(&raw);
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = static_cast<float>(h[j]) * s;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// the pass's output plane: streaming (non-temporal) 16-byte store
template <typename T>
__device__ __forceinline__ void store8(void *plane, size_t p, float inv, const float (&v)[8]) {
    u32x4 raw;
    T *h = reinterpret_cast<T *>(&raw);
#pragma unroll
    for (int j = 0; j < 8; ++j) h[j] = static_cast<T>(v[j] * inv);
    __builtin_nontemporal_store(raw, reinterpret_cast<u32x4 *>(static_cast<T *>(plane) + p));
}

// the plane a pass reads exactly once: streaming load
template <typename T>
__device__ __forceinline__ void load8s(const void *plane, size_t p, float s, float (&o)[8]) {
    const u32x4 raw = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(static_cast<const T *>(plane) + p));
    const T *h = reinterpret_cast<const T *>(&raw);
#pragma unroll
    for (int j = 0; j < 8; ++j) o[j] = static_cast<float>(h[j]) * s;
}

template <typename T>
__device__ __forceinline__ float load1(const void *plane, size_t p, float s) {
    return static_cast<float>(static_cast<const T *>(plane)[p]) * s;
}

__device__ __forceinline__ void load_ids8(const LatticeArgs &a, size_t p, uint32_t (&id)[8]) {
    if (!a.mat_id) {
#pragma unroll
        for (int j = 0; j < 8; ++j) id[j] = 0u;
    } else if (a.id_bytes == 2) {
        const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(a.mat_id) + p);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            id[2 * j] = w[j] & 0xffffu;
            id[2 * j + 1] = w[j] >> 16;
        }
    } else {
        const uint4 v0 = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(a.mat_id) + p);
        const uint4 v1 = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(a.mat_id) + p + 4);
        id[0] = v0.x; id[1] = v0.y; id[2] = v0.z; id[3] = v0.w;
        id[4] = v1.x; id[5] = v1.y; id[6] = v1.z; id[7] = v1.w;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_update_b8h(LatticeArgs a, FieldPtrs f) {
    const int x0 = (blockIdx.x * 32 + threadIdx.x) * 8, y = blockIdx.y * 8 + threadIdx.y, z = blockIdx.z;
    if (x0 >= a.nx || y >= a.ny) return;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x0;
    const bool up_x = x0 + 8 < a.nx, up_y = y + 1 < a.ny, up_z = z + 1 < a.nz;
    float e0[3][8], eX[3], eY[3][8], eZ[3][8], b[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float s = f.scale[c];
        load8<T>(f.e[c], p, s, e0[c]);
        eX[c] = up_x ? load1<T>(f.e[c], p + 8, s) : 0.0f;
        if (up_y) load8<T>(f.e[c], p + sy, s, eY[c]);
        else
#pragma unroll
            for (int j = 0; j < 8; ++j) eY[c][j] = 0.0f;
        if (up_z) load8<T>(f.e[c], p + sz, s, eZ[c]);
        else
#pragma unroll
            for (int j = 0; j < 8; ++j) eZ[c][j] = 0.0f;
        load8s<T>(f.b[c], p, f.scale[3 + c], b[c]);
    }
    // a class only for thin-wire factors, and only where the wire mask marks one of the 8 texels
    uint32_t id[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
    if (a.has_wire) {
        if (a.wire_mask) {
            const uint32_t w = a.wire_mask[(static_cast<size_t>(z) * a.ny + y) * a.mask_wpr + (x0 >> 5)];
            if ((w >> (x0 & 31)) & 0xffu) load_ids8(a, p, id);
        } else {
            load_ids8(a, p, id);
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float c0[3] = {e0[0][j], e0[1][j], e0[2][j]};
        const float cx[3] = {j < 7 ? e0[0][j + 1] : eX[0], j < 7 ? e0[1][j + 1] : eX[1], j < 7 ? e0[2][j + 1] : eX[2]};
        const float cy[3] = {eY[0][j], eY[1][j], eY[2][j]};
        const float cz[3] = {eZ[0][j], eZ[1][j], eZ[2][j]};
        float bb[3] = {b[0][j], b[1][j], b[2][j]};
        faraday(a, x0 + j, y, z, c0, cx, cy, cz, id[j], bb);
        b[0][j] = bb[0];
        b[1][j] = bb[1];
        b[2][j] = bb[2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store8<T>(f.b[c], p, 1.0f / f.scale[3 + c], b[c]);
}

// PM: see edge_coefficients (false only for lattices without a material texel in the absorber)
template <typename T, bool PM>
__global__ void __launch_bounds__(256) k_update_e8h(LatticeArgs a, FieldPtrs f) {
    const int x0 = (blockIdx.x * 32 + threadIdx.x) * 8, y = blockIdx.y * 8 + threadIdx.y, z = blockIdx.z;
    if (x0 >= a.nx || y >= a.ny) return;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x0;
    const bool dn_x = x0 > 0, dn_y = y > 0, dn_z = z > 0;
    float b0[3][8], bX[3], bY[3][8], bZ[3][8], e[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float s = f.scale[3 + c];
        load8<T>(f.b[c], p, s, b0[c]);
        bX[c] = dn_x ? load1<T>(f.b[c], p - 1, s) : 0.0f;
        if (dn_y) load8<T>(f.b[c], p - sy, s, bY[c]);
        else
#pragma unroll
            for (int j = 0; j < 8; ++j) bY[c][j] = 0.0f;
        if (dn_z) load8<T>(f.b[c], p - sz, s, bZ[c]);
        else
#pragma unroll
            for (int j = 0; j < 8; ++j) bZ[c][j] = 0.0f;
        load8s<T>(f.e[c], p, f.scale[c], e[c]);
    }
    // (a class-mask gate as in the FP32 pass measured slower here: 71.8 -> 80.5 us at 256^3; the
    // mask-then-class round trip stalls the waves that hold a material texel)
    uint32_t id[8];
    load_ids8(a, p, id);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const float bxm[3] = {j > 0 ? b0[0][j - 1] : bX[0], j > 0 ? b0[1][j - 1] : bX[1], j > 0 ? b0[2][j - 1] : bX[2]};
        const float curl[3] = {(b0[2][j] - bY[2][j]) - (b0[1][j] - bZ[1][j]), (b0[0][j] - bZ[0][j]) - (b0[2][j] - bxm[2]),
                               (b0[1][j] - bxm[1]) - (b0[0][j] - bY[0][j])};
        float ca[3], cb[3];
        edge_coefficients<PM>(a, x0 + j, y, z, id[j], ca, cb);
#pragma unroll
        for (int c = 0; c < 3; ++c) e[c][j] = ca[c] * e[c][j] + cb[c] * curl[c];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) store8<T>(f.e[c], p, 1.0f / f.scale[c], e[c]);
}

template <typename T>
hipError_t launch_t(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s) {
    dim3 block(32, 8, 1), grid((a.nx / 8 + 31) / 32, (a.ny + 7) / 8, a.nz);
    k_update_b8h<T><<<grid, block, 0, s>>>(a, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.pml_materials)
        k_update_e8h<T, true><<<grid, block, 0, s>>>(a, f);
    else
        k_update_e8h<T, false><<<grid, block, 0, s>>>(a, f);
    return hipGetLastError();
}

}  // namespace

int yee_half_supported(const LatticeArgs &a) { return a.nx % 8 == 0 ? 1 : 0; }

hipError_t launch_yee_half(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s) {
    if (!yee_half_supported(a)) return hipErrorInvalidValue;
    if (f.storage == kStorageF16) return launch_t<_Float16>(a, f, s);
    if (f.storage == kStorageBF16) return launch_t<__bf16>(a, f, s);
    return hipErrorInvalidValue;
}

}  // namespace emfd
