// Two-pass leapfrog on 16-bit field storage: E and B planes hold binary16 or bfloat16 values of
// field / scale (a power-of-two scale per component), every update runs in FP32 with the same
// per-texel arithmetic as faraday() / edge_coefficients(), and the result is rounded to nearest
// even on the way back. Each thread owns N = 8 (16-byte loads and stores) or N = 4 (8-byte)
// consecutive x cells. Requires nx % 8 == 0. Field traffic per cell-update is 24 B (half of
// FP32), plus the x+-1 neighbours of the edge cells, which hit the same lines.
#include "coef_rows.h"

namespace emfd {
namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <int N>
struct Raw;  // N 16-bit values
template <>
struct Raw<8> {
    typedef u32x4 type;
};
template <>
struct Raw<4> {
    typedef u32x2 type;
};

// N values of a plane; NT: the plane a pass reads exactly once (streaming load)
template <typename T, int N, bool NT>
__device__ __forceinline__ void loadn(const void *plane, size_t p, float s, float (&o)[N]) {
    using R = typename Raw<N>::type;
    const R *ptr = reinterpret_cast<const R *>(static_cast<const T *>(plane) + p);
    const R raw = NT ? __builtin_nontemporal_load(ptr) : *ptr;
    const T *h = reinterpret_cast<const T *>(&raw);
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = static_cast<float>(h[j]) * s;
}

// the pass's output plane: streaming (non-temporal) store
template <typename T, int N>
__device__ __forceinline__ void storen(void *plane, size_t p, float inv, const float (&v)[N]) {
    using R = typename Raw<N>::type;
    R raw;
    T *h = reinterpret_cast<T *>(&raw);
#pragma unroll
    for (int j = 0; j < N; ++j) h[j] = static_cast<T>(v[j] * inv);
    __builtin_nontemporal_store(raw, reinterpret_cast<R *>(static_cast<T *>(plane) + p));
}

// loadn() of a neighbour that may lie beyond the lattice: zeros (+0, as the reference's zero
// boundary) when !in. The load is unconditional, at `own` (the cell's own, always valid, index)
// when !in, and the raw bits are selected after it: a load inside a branch was converted inside
// it too, so the wave waited on each such load in turn instead of once for all.
template <typename T, int N>
__device__ __forceinline__ void loadn_or_zero(const void *plane, size_t p, size_t own, bool in, float s,
                                              float (&o)[N]) {
    using R = typename Raw<N>::type;
    R raw = *reinterpret_cast<const R *>(static_cast<const T *>(plane) + (in ? p : own));
#pragma unroll
    for (int k = 0; k < N / 2; ++k) raw[k] = in ? raw[k] : 0u;
    const T *h = reinterpret_cast<const T *>(&raw);
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = static_cast<float>(h[j]) * s;
}

// loadn_or_zero() without the conversion: the N stored values themselves (+0 when !in)
template <typename T, int N>
__device__ __forceinline__ void loadn_raw_or_zero(const void *plane, size_t p, size_t own, bool in, T (&o)[N]) {
    using R = typename Raw<N>::type;
    R raw = *reinterpret_cast<const R *>(static_cast<const T *>(plane) + (in ? p : own));
#pragma unroll
    for (int k = 0; k < N / 2; ++k) raw[k] = in ? raw[k] : 0u;
    const T *h = reinterpret_cast<const T *>(&raw);
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = h[j];
}

// c - h * s as the separate operations round it (h * s is exact: s is a power of two), in one
// mixed-precision fused multiply-add
template <typename T>
__device__ __forceinline__ float minus_scaled(float c, T h, float s) {
    return __builtin_fmaf(-static_cast<float>(h), s, c);
}

// h * s - c as the separate operations round it, in one mixed-precision fused multiply-add
template <typename T>
__device__ __forceinline__ float scaled_minus(T h, float s, float c) {
    return __builtin_fmaf(static_cast<float>(h), s, -c);
}

template <typename T>
__device__ __forceinline__ float load1_or_zero(const void *plane, size_t p, size_t own, bool in, float s) {
    const T v = static_cast<const T *>(plane)[in ? p : own];
    return in ? static_cast<float>(v) * s : 0.0f;
}

// Copies the stored values of components comp0..comp0+2 (rounded through T, times the scale: what
// a reload returns) of the N cells into the far-field side columns when the run holds one of them.
template <typename T, int N>
__device__ __forceinline__ void side_write_h(const LatticeArgs &a, const FieldPtrs &f, int comp0, int x0, int y, int z,
                                             const float (&v)[3][N]) {
    side_write_cells<N>(a.side, comp0, x0, y, z, [&](int c, int j) {
        return static_cast<float>(static_cast<T>(v[c][j] * f.inv_scale[comp0 + c])) * f.scale[comp0 + c];
    });
}

// load_idsn() with the id width fixed at compile time (IDB: 0 no class ids, 2 or 4 bytes): a
// run-time width branch in front of the load made the compiler wait for every field load issued
// before it, a second round trip per wavefront.
template <int N, int IDB>
__device__ __forceinline__ void load_idsn_fixed(const LatticeArgs &a, size_t p, uint32_t (&id)[N]) {
    if (IDB == 0) {
#pragma unroll
        for (int j = 0; j < N; ++j) id[j] = 0u;
    } else if (IDB == 2) {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const uint2 v = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(a.mat_id) + p + 4 * q);
            id[4 * q + 0] = v.x & 0xffffu;
            id[4 * q + 1] = v.x >> 16;
            id[4 * q + 2] = v.y & 0xffffu;
            id[4 * q + 3] = v.y >> 16;
        }
    } else {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(a.mat_id) + p + 4 * q);
            id[4 * q + 0] = v.x;
            id[4 * q + 1] = v.y;
            id[4 * q + 2] = v.z;
            id[4 * q + 3] = v.w;
        }
    }
}

template <int N>
__device__ __forceinline__ void load_idsn(const LatticeArgs &a, size_t p, uint32_t (&id)[N]) {
    if (!a.mat_id) {
#pragma unroll
        for (int j = 0; j < N; ++j) id[j] = 0u;
    } else if (a.id_bytes == 2) {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const uint2 v = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(a.mat_id) + p + 4 * q);
            id[4 * q + 0] = v.x & 0xffffu;
            id[4 * q + 1] = v.x >> 16;
            id[4 * q + 2] = v.y & 0xffffu;
            id[4 * q + 3] = v.y >> 16;
        }
    } else {
#pragma unroll
        for (int q = 0; q < N / 4; ++q) {
            const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(a.mat_id) + p + 4 * q);
            id[4 * q + 0] = v.x;
            id[4 * q + 1] = v.y;
            id[4 * q + 2] = v.z;
            id[4 * q + 3] = v.w;
        }
    }
}

// workgroup shape: TX threads along x (TX * N cells), 256 / TX rows; a wavefront holds kRw<N> rows
template <int N>
constexpr int kTx = N == 8 ? 32 : 64;
template <int N>
constexpr int kRw = 64 / kTx<N>;

template <typename T, int N>
__global__ void __launch_bounds__(256) k_update_bh(LatticeArgs a, FieldPtrs f) {
    extern __shared__ float coef_lds[];
    const int x0 = (blockIdx.x * kTx<N> + threadIdx.x) * N, y = blockIdx.y * (256 / kTx<N>) + threadIdx.y,
              z = blockIdx.z;
    const int wave = threadIdx.y / kRw<N>, lane = (threadIdx.y % kRw<N>) * kTx<N> + threadIdx.x;
    const int y_first = blockIdx.y * (256 / kTx<N>) + wave * kRw<N>;
    float *rows = coef_lds + wave * coef_entries(a, kRw<N>) * kCoefStride;
    const CoefLoad coef0 = coef_load<kRw<N>>(a, lane, y_first, z, kPmlBa);
    // lanes beyond the lattice load at cell (0, 0, z) and return after the staging
    const bool active = x0 < a.nx && y < a.ny;
    const int xa = active ? x0 : 0, ya = active ? y : 0;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + ya * sy + xa;
    const bool up_x = xa + N < a.nx, up_y = ya + 1 < a.ny, up_z = z + 1 < a.nz;
    // the +y and +z neighbours enter the curl once each, as neighbour - e0: kept as stored values
    // and folded into scaled_minus(); the +y neighbour's y component and the +z neighbour's z
    // component are not used
    float e0[3][N], eX[3], b[3][N];
    T hY[3][N], hZ[3][N];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float s = f.scale[c];
        loadn<T, N, false>(f.e[c], p, s, e0[c]);
        eX[c] = load1_or_zero<T>(f.e[c], p + N, p, up_x, s);
        if (c != 1) loadn_raw_or_zero<T, N>(f.e[c], p + sy, p, up_y, hY[c]);
        if (c != 2) loadn_raw_or_zero<T, N>(f.e[c], p + sz, p, up_z, hZ[c]);
        loadn<T, N, true>(f.b[c], p, f.scale[3 + c], b[c]);
    }
    const float sE[3] = {f.scale[0], f.scale[1], f.scale[2]};
    // a class only for thin-wire factors, and only where the wire mask marks one of the N texels
    uint32_t id[N];
#pragma unroll
    for (int j = 0; j < N; ++j) id[j] = 0u;
    if (a.has_wire) {
        if (a.wire_mask) {
            const uint32_t w = a.wire_mask[(static_cast<size_t>(z) * a.ny + ya) * a.mask_wpr + (xa >> 5)];
            if ((w >> (xa & 31)) & ((1u << N) - 1u)) load_idsn<N>(a, p, id);
        } else {
            load_idsn<N>(a, p, id);
        }
    }
    coef_store(a, rows, lane, coef0);
    coef_finish<kRw<N>>(a, rows, lane, y_first, z, kPmlBa);
    if (!active) return;
    const float *row = rows + (threadIdx.y % kRw<N>) * a.pml_k * kCoefStride;
    // thin-wire texels (rare: ids are loaded only where the wire mask marks one) take faraday()
    // with its per-texel branches; every other lane one straight path with staged coefficients
    bool wired = false;
#pragma unroll
    for (int j = 0; j < N; ++j) wired |= id[j] != 0u;
    if (!wired) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            // faraday_coef() without wire factors: the same terms, in the same order
            const float xn1 = j < N - 1 ? e0[1][j + 1] : eX[1], xn2 = j < N - 1 ? e0[2][j + 1] : eX[2];
            const float t0x = scaled_minus(hY[2][j], sE[2], e0[2][j]), t0y = scaled_minus(hZ[0][j], sE[0], e0[0][j]),
                        t0z = xn1 - e0[1][j];
            const float t1x = scaled_minus(hZ[1][j], sE[1], e0[1][j]), t1y = xn2 - e0[2][j],
                        t1z = scaled_minus(hY[0][j], sE[0], e0[0][j]);
            const float cx = t0x - t1x, cy = t0y - t1y, cz = t0z - t1z;
            float ka[3], kb[3];
            read_coef(row + pml_class(x0 + j, a.nx, a.pml) * kCoefStride, ka, kb);
            b[0][j] = ka[0] * b[0][j] - kb[0] * cx;
            b[1][j] = ka[1] * b[1][j] - kb[1] * cy;
            b[2][j] = ka[2] * b[2][j] - kb[2] * cz;
        }
    } else {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const float c0[3] = {e0[0][j], e0[1][j], e0[2][j]};
            const float cx[3] = {j < N - 1 ? e0[0][j + 1] : eX[0], j < N - 1 ? e0[1][j + 1] : eX[1],
                                 j < N - 1 ? e0[2][j + 1] : eX[2]};
            const float cy[3] = {static_cast<float>(hY[0][j]) * sE[0], 0.0f, static_cast<float>(hY[2][j]) * sE[2]};
            const float cz[3] = {static_cast<float>(hZ[0][j]) * sE[0], static_cast<float>(hZ[1][j]) * sE[1], 0.0f};
            float bb[3] = {b[0][j], b[1][j], b[2][j]};
            faraday(a, x0 + j, y, z, c0, cx, cy, cz, id[j], bb);
            b[0][j] = bb[0];
            b[1][j] = bb[1];
            b[2][j] = bb[2];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) storen<T, N>(f.b[c], p, f.inv_scale[3 + c], b[c]);
    side_write_h<T, N>(a, f, 3, x0, y, z, b);
}

// PM: see edge_coefficients (false only for lattices without a material texel in the absorber);
// IDB: class id width (0: no ids)
template <typename T, int N, bool PM, int IDB>
__global__ void __launch_bounds__(256) k_update_eh(LatticeArgs a, FieldPtrs f) {
    extern __shared__ float coef_lds[];
    const int x0 = (blockIdx.x * kTx<N> + threadIdx.x) * N, y = blockIdx.y * (256 / kTx<N>) + threadIdx.y,
              z = blockIdx.z;
    const int wave = threadIdx.y / kRw<N>, lane = (threadIdx.y % kRw<N>) * kTx<N> + threadIdx.x;
    const int y_first = blockIdx.y * (256 / kTx<N>) + wave * kRw<N>;
    float *rows = coef_lds + wave * coef_entries(a, kRw<N>) * kCoefStride;
    const CoefLoad coef0 = coef_load<kRw<N>>(a, lane, y_first, z, kPmlEa);
    // lanes beyond the lattice load at cell (0, 0, z) and return after the staging
    const bool active = x0 < a.nx && y < a.ny;
    const int xa = active ? x0 : 0, ya = active ? y : 0;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + ya * sy + xa;
    const bool dn_x = xa > 0, dn_y = ya > 0, dn_z = z > 0;
    // the -y and -z neighbours enter the curl once each, as b0 - neighbour: kept as stored values
    // and folded into minus_scaled(); the -y neighbour's y component and the -z neighbour's z
    // component are not used
    float b0[3][N], bX[3], e[3][N];
    T hY[3][N], hZ[3][N];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float s = f.scale[3 + c];
        loadn<T, N, false>(f.b[c], p, s, b0[c]);
        bX[c] = load1_or_zero<T>(f.b[c], p - 1, p, dn_x, s);
        if (c != 1) loadn_raw_or_zero<T, N>(f.b[c], p - sy, p, dn_y, hY[c]);
        if (c != 2) loadn_raw_or_zero<T, N>(f.b[c], p - sz, p, dn_z, hZ[c]);
        loadn<T, N, true>(f.e[c], p, f.scale[c], e[c]);
    }
    const float sB[3] = {f.scale[3], f.scale[4], f.scale[5]};
    // (a class-mask gate as in the FP32 pass measured slower here: 71.8 -> 80.5 us at 256^3; the
    // mask-then-class round trip stalls the waves that hold a material texel)
    uint32_t id[N];
    load_idsn_fixed<N, IDB>(a, p, id);
    coef_store(a, rows, lane, coef0);
    coef_finish<kRw<N>>(a, rows, lane, y_first, z, kPmlEa);
    if (!active) return;
    const float *row = rows + (threadIdx.y % kRw<N>) * a.pml_k * kCoefStride;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const float bxm[3] = {j > 0 ? b0[0][j - 1] : bX[0], j > 0 ? b0[1][j - 1] : bX[1], j > 0 ? b0[2][j - 1] : bX[2]};
        const float curl[3] = {minus_scaled(b0[2][j], hY[2][j], sB[2]) - minus_scaled(b0[1][j], hZ[1][j], sB[1]),
                               minus_scaled(b0[0][j], hZ[0][j], sB[0]) - (b0[2][j] - bxm[2]),
                               (b0[1][j] - bxm[1]) - minus_scaled(b0[0][j], hY[0][j], sB[0])};
        // edge_coefficients() with the vacuum edges' (absorber or interior) coefficients staged
        float ca[3], cb[3];
        // every texel reads its staged entry; a material texel then replaces it (read first and
        // unconditionally, the vacuum entry does not wait behind the material texels' class loads)
        read_coef(row + pml_class(x0 + j, a.nx, a.pml) * kCoefStride, ca, cb);
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
    for (int c = 0; c < 3; ++c) storen<T, N>(f.e[c], p, f.inv_scale[c], e[c]);
    side_write_h<T, N>(a, f, 0, x0, y, z, e);
}

template <typename T, int N>
hipError_t launch_t(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s) {
    constexpr int tx = kTx<N>, ty = 256 / kTx<N>;
    dim3 block(tx, ty, 1), grid((a.nx / N + tx - 1) / tx, (a.ny + ty - 1) / ty, a.nz);
    // staged absorber rows: 4 wavefronts x coef_entries (pml <= 64: at most 33 KB)
    const size_t lds = static_cast<size_t>(4 * coef_entries(a, kRw<N>) * kCoefStride) * sizeof(float);
    k_update_bh<T, N><<<grid, block, lds, s>>>(a, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int idb = a.mat_id ? a.id_bytes : 0;
    if (a.pml_materials) {
        if (idb == 2) k_update_eh<T, N, true, 2><<<grid, block, lds, s>>>(a, f);
        else if (idb == 4) k_update_eh<T, N, true, 4><<<grid, block, lds, s>>>(a, f);
        else k_update_eh<T, N, true, 0><<<grid, block, lds, s>>>(a, f);
    } else {
        if (idb == 2) k_update_eh<T, N, false, 2><<<grid, block, lds, s>>>(a, f);
        else if (idb == 4) k_update_eh<T, N, false, 4><<<grid, block, lds, s>>>(a, f);
        else k_update_eh<T, N, false, 0><<<grid, block, lds, s>>>(a, f);
    }
    return hipGetLastError();
}

}  // namespace

int yee_half_supported(const LatticeArgs &a) { return a.nx % 8 == 0 ? 1 : 0; }

// cells_per_thread: 8 or 4 (0: 8)
hipError_t launch_yee_half(const LatticeArgs &a, const FieldPtrs &f, int cells_per_thread, hipStream_t s) {
    if (!yee_half_supported(a)) return hipErrorInvalidValue;
    const bool four = cells_per_thread == 4;
    if (cells_per_thread != 0 && cells_per_thread != 4 && cells_per_thread != 8) return hipErrorInvalidValue;
    if (f.storage == kStorageF16) return four ? launch_t<_Float16, 4>(a, f, s) : launch_t<_Float16, 8>(a, f, s);
    if (f.storage == kStorageBF16) return four ? launch_t<__bf16, 4>(a, f, s) : launch_t<__bf16, 8>(a, f, s);
    return hipErrorInvalidValue;
}

}  // namespace emfd
