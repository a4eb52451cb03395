// HIP kernels of the FDTD hot path for gfx950 (MI355X).
//
// Field storage: six float planes (Ex Ey Ez Bx By Bz), each x-fastest with EM_Lab's lattice index
// (z*ny + y)*nx + x, so a 64-lane wavefront reads 256 contiguous bytes per component. The cell's
// material enters through a 2-byte class id that indexes a small coefficient table (absent for
// uniform vacuum); the absorber conductivity is rebuilt from three 1-D profiles.
//
// Arithmetic follows the reference shaders operation by operation (shaders/common.glsl:21-54,
// shaders/update_B.comp:31-70, shaders/update_E.comp:31-66, shaders/apply_sources.comp:37-79);
// this file is compiled with -ffp-contract=off so no multiply-add is fused, and every exponential
// goes through det_expf, which gives the same bits on the host and the device.
#include "expf.h"
#include "internal.h"
#include "sources_device.h"
#include "yee_device.h"

#include <algorithm>

namespace emfd {
namespace {


// ---------------------------------------------------------------------------------------------
// Coefficient table from binary16 keys: 12 halves per class
// (sigma x y z, eps x y z, wire Bx t0 t1, By t0 t1, Bz t0 t1).
// ---------------------------------------------------------------------------------------------
__global__ void k_build_table(const uint16_t *keys, int classes, float dt, float *table) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= classes) return;
    const uint16_t *k = keys + 12 * static_cast<size_t>(i);
    float *t = table + static_cast<size_t>(i) * kTableStride;
    float h[12];
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        _Float16 v;
        __builtin_memcpy(&v, &k[j], 2);
        h[j] = static_cast<float>(v);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float sig = fmaxf(h[c], 0.0f);
        const float eps = fmaxf(h[3 + c], 1.0f);
        const float dtEps = dt / eps;
        const float xx = sig * dtEps;
        t[kSig + c] = sig;
        t[kEps + c] = eps;
        t[kCa + c] = det_expf(-xx);
        t[kCb + c] = dtEps * phi_f(xx);
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) t[kWire + j] = h[6 + j];
    t[18] = 0.0f;
    t[19] = 0.0f;
}

// Absorber coefficient table (one entry per (kx, ky, kz) depth class), built with the same
// expressions faraday() / edge_coefficients() would evaluate per texel.
__device__ __forceinline__ int pml_index_of(int k, int n, int L) {
    return k == 0 ? L : (k <= L ? k - 1 : (n - 1 - L) + (k - 1 - L));
}

__global__ void k_build_pml_table(LatticeArgs a, float *tab) {
    const int K = a.pml_k;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K * K * K) return;
    const int x = pml_index_of(e % K, a.nx, a.pml), y = pml_index_of((e / K) % K, a.ny, a.pml),
              z = pml_index_of(e / (K * K), a.nz, a.pml);
    const float dt = a.dt;
    float *t = tab + static_cast<size_t>(e) * kPmlStride;
    const float sB[3] = {pml_sigma(a, x, 0, y, 1, z, 1) * dt, pml_sigma(a, x, 1, y, 0, z, 1) * dt,
                         pml_sigma(a, x, 1, y, 1, z, 0) * dt};
    const float sE[3] = {0.0f + pml_sigma(a, x, 1, y, 0, z, 0), 0.0f + pml_sigma(a, x, 0, y, 1, z, 0),
                         0.0f + pml_sigma(a, x, 0, y, 0, z, 1)};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        t[kPmlBa + c] = det_expf(-sB[c]);
        t[kPmlBb + c] = dt * phi_f(sB[c]);
        const float dtEps = dt / 1.0f;
        const float xx = sE[c] * dtEps;
        t[kPmlEa + c] = det_expf(-xx);
        t[kPmlEb + c] = dtEps * phi_f(xx);
    }
    t[12] = t[13] = t[14] = t[15] = 0.0f;
    if (e == 0) {
        // no absorber on any axis: the interior coefficients, read by kernels that load every
        // texel's entry (1 * b - dt * c is b - dt * c bit for bit; (1, dt) are the vacuum edge's)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            t[kPmlBa + c] = 1.0f;
            t[kPmlBb + c] = dt;
            t[kPmlEa + c] = 1.0f;
            t[kPmlEb + c] = dt;
        }
    }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Plain float4 copy: the bandwidth and counter calibration reference for the roofline figures.
__global__ void __launch_bounds__(256) k_copy4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

// Stream probe: R planes read and W planes written at the same index, two float4 of each plane
// per thread (every load issued before any store), non-temporal stores; the access pattern of the
// fused Yee pass (R = W = 6) with the stencil taken out. n: float4 per plane, a multiple of
// 256 * 2 * gridDim.x.
template <int R, int W>
__global__ void __launch_bounds__(256) k_stream(const f32x4 *__restrict__ a, f32x4 *__restrict__ b, size_t n,
                                                size_t plane) {
    constexpr int U = 2;
    const size_t stride = static_cast<size_t>(gridDim.x) * 256 * U;
    for (size_t base = static_cast<size_t>(blockIdx.x) * 256 * U + threadIdx.x; base < n; base += stride) {
        f32x4 acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) acc[u] = a[base + u * 256];
#pragma unroll
        for (int r = 1; r < R; ++r)
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u] += a[r * plane + base + u * 256];
#pragma unroll
        for (int w = 0; w < W; ++w)
#pragma unroll
            for (int u = 0; u < U; ++u)
                __builtin_nontemporal_store(acc[u] * static_cast<float>(w + 1), &b[w * plane + base + u * 256]);
    }
}

__global__ void k_fill(float *p, size_t n, float v) {
    size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------------------------------------
// Two-pass variant: update_B then update_E in place, one thread per texel, as the reference
// dispatches them (src/SimulationGrid.cpp:393-408). Kept as a plain baseline.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_update_b(LatticeArgs a, FieldPtrs f) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= a.nx || y >= a.ny) return;
    const size_t sx = 1, sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x;
    float e0[3], exn[3], eyn[3], ezn[3], b[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *e = f.e[c];
        e0[c] = e[p];
        exn[c] = x + 1 < a.nx ? e[p + sx] : 0.0f;
        eyn[c] = y + 1 < a.ny ? e[p + sy] : 0.0f;
        ezn[c] = z + 1 < a.nz ? e[p + sz] : 0.0f;
        b[c] = f.b[c][p];
    }
    const uint32_t id = a.mat_id ? load_id(a, p) : 0u;
    faraday(a, x, y, z, e0, exn, eyn, ezn, id, b);
#pragma unroll
    for (int c = 0; c < 3; ++c) f.b[c][p] = b[c];
}

__global__ void __launch_bounds__(256) k_update_e(LatticeArgs a, FieldPtrs f) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= a.nx || y >= a.ny) return;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x;
    float b0[3], bxm[3], bym[3], bzm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *b = f.b[c];
        b0[c] = b[p];
        bxm[c] = x > 0 ? b[p - 1] : 0.0f;
        bym[c] = y > 0 ? b[p - sy] : 0.0f;
        bzm[c] = z > 0 ? b[p - sz] : 0.0f;
    }
    const float curl[3] = {(b0[2] - bym[2]) - (b0[1] - bzm[1]), (b0[0] - bzm[0]) - (b0[2] - bxm[2]),
                           (b0[1] - bxm[1]) - (b0[0] - bym[0])};
    const uint32_t id = a.mat_id ? load_id(a, p) : 0u;
    float ca[3], cb[3];
    edge_coefficients(a, x, y, z, id, ca, cb);
#pragma unroll
    for (int c = 0; c < 3; ++c) f.e[c][p] = ca[c] * f.e[c][p] + cb[c] * curl[c];
}

// Per-texel class masks, bit x % 32 of word (z ny + y) wpr + x / 32: wire_mask marks texels whose
// class carries a thin-wire factor other than 1 (only those need their class in the Faraday
// update: a factor of exactly 1 leaves every product unchanged); class_mask marks texels whose
// class is not 0 (the Ampere update of a class-0 texel needs no class record).
__global__ void __launch_bounds__(256) k_build_class_masks(LatticeArgs a, uint32_t *wire_mask, uint32_t *class_mask,
                                                           int wpr) {
    const size_t w = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t rows = static_cast<size_t>(a.ny) * a.nz;
    if (w >= rows * wpr) return;
    const size_t row = w / wpr;
    const int x0 = static_cast<int>(w % wpr) * 32;
    uint32_t wires = 0u, classes = 0u;
    for (int b = 0; b < 32 && x0 + b < a.nx; ++b) {
        const uint32_t id = load_id(a, row * a.nx + x0 + b);
        if (id == 0u) continue;
        classes |= 1u << b;
        const float *t = a.table + static_cast<size_t>(id) * kTableStride + kWire;
        bool wire = false;
        for (int k = 0; k < 6; ++k) wire = wire || t[k] != 1.0f;
        if (wire) wires |= 1u << b;
    }
    wire_mask[w] = wires;
    class_mask[w] = classes;
}

__global__ void __launch_bounds__(64) k_apply_sources(LatticeArgs a, FieldPtrs f, const int *cells, int n_cells,
                                                      const int *lists, const float *packed, const float *signals) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_cells) return;
    source_cell(a, f, cells, lists, packed, signals, id);
}

__global__ void __launch_bounds__(64) k_build_source_records(LatticeArgs a, const int *cells, int n_cells,
                                                              const int *lists, const float *packed, SourceRec *recs,
                                                              SourcePair *pairs) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_cells) return;
    source_record(a, cells, lists, packed, id, recs, pairs);
}

// S: the field storage, fixed at compile time so the field loads carry no format branch
template <int S>
__global__ void __launch_bounds__(64) k_apply_source_records(LatticeArgs a, FieldPtrs f, const SourceRec *recs,
                                                             const SourcePair *pairs, int n_cells,
                                                             const float *packed, const float *signals) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_cells) return;
    f.storage = S;
    source_apply(a, f, recs, pairs, packed, signals, id);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------
hipError_t launch_build_table(const uint16_t *keys, int classes, float dt, float *table, hipStream_t s) {
    if (classes <= 0) return hipSuccess;
    k_build_table<<<(classes + 127) / 128, 128, 0, s>>>(keys, classes, dt, table);
    return hipGetLastError();
}

hipError_t launch_build_class_masks(const LatticeArgs &a, uint32_t *wire_mask, uint32_t *class_mask, int wpr,
                                   hipStream_t s) {
    const size_t words = static_cast<size_t>(a.ny) * a.nz * wpr;
    if (words == 0 || !a.mat_id) return hipErrorInvalidValue;
    k_build_class_masks<<<static_cast<unsigned>((words + 255) / 256), 256, 0, s>>>(a, wire_mask, class_mask, wpr);
    return hipGetLastError();
}

hipError_t launch_build_pml_table(const LatticeArgs &a, float *tab, hipStream_t s) {
    const int n = a.pml_k * a.pml_k * a.pml_k;
    k_build_pml_table<<<(n + 127) / 128, 128, 0, s>>>(a, tab);
    return hipGetLastError();
}

hipError_t launch_copy4(const void *src, void *dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    k_copy4<<<16384, 256, 0, s>>>(static_cast<const float4 *>(src), static_cast<float4 *>(dst), n);
    return hipGetLastError();
}

hipError_t launch_stream_probe(const void *src, void *dst, size_t plane_f4, int planes, hipStream_t s) {
    // plane_f4 is a multiple of 256 * 2 * 4096 (the caller rounds it)
    const unsigned blocks = 4096;
    const f32x4 *a = static_cast<const f32x4 *>(src);
    f32x4 *b = static_cast<f32x4 *>(dst);
    switch (planes) {
    case 1: k_stream<1, 1><<<blocks, 256, 0, s>>>(a, b, plane_f4, plane_f4); break;
    case 2: k_stream<2, 2><<<blocks, 256, 0, s>>>(a, b, plane_f4, plane_f4); break;
    case 3: k_stream<3, 3><<<blocks, 256, 0, s>>>(a, b, plane_f4, plane_f4); break;
    default: k_stream<6, 6><<<blocks, 256, 0, s>>>(a, b, plane_f4, plane_f4); break;
    }
    return hipGetLastError();
}

hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 65536);
    k_fill<<<static_cast<unsigned>(blocks), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_yee_two_pass(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s, int passes) {
    dim3 block(64, 4, 1), grid((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
    if (passes & 1) {
        k_update_b<<<grid, block, 0, s>>>(a, f);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (passes & 2) k_update_e<<<grid, block, 0, s>>>(a, f);
    return hipGetLastError();
}

hipError_t launch_apply_sources(const LatticeArgs &a, const FieldPtrs &f, const int *cells, int n_cells,
                                const int *lists, const float *packed, const float *signals, hipStream_t s) {
    if (n_cells <= 0) return hipSuccess;
    k_apply_sources<<<(n_cells + 63) / 64, 64, 0, s>>>(a, f, cells, n_cells, lists, packed, signals);
    return hipGetLastError();
}

hipError_t launch_build_source_records(const LatticeArgs &a, const int *cells, int n_cells, const int *lists,
                                       const float *packed, void *recs, void *pairs, hipStream_t s) {
    static_assert(sizeof(SourceRec) == kSourceRecBytes && sizeof(SourcePair) == kSourcePairBytes, "record layout");
    if (n_cells <= 0) return hipSuccess;
    k_build_source_records<<<(n_cells + 63) / 64, 64, 0, s>>>(a, cells, n_cells, lists, packed,
                                                              static_cast<SourceRec *>(recs),
                                                              static_cast<SourcePair *>(pairs));
    return hipGetLastError();
}

hipError_t launch_apply_source_records(const LatticeArgs &a, const FieldPtrs &f, const void *recs, const void *pairs,
                                       int n_cells, const float *packed, const float *signals, hipStream_t s) {
    if (n_cells <= 0) return hipSuccess;
    const dim3 g((n_cells + 63) / 64);
    const SourceRec *r = static_cast<const SourceRec *>(recs);
    const SourcePair *q = static_cast<const SourcePair *>(pairs);
    switch (f.storage) {
        case kStorageF32: k_apply_source_records<kStorageF32><<<g, 64, 0, s>>>(a, f, r, q, n_cells, packed, signals); break;
        case kStorageF16: k_apply_source_records<kStorageF16><<<g, 64, 0, s>>>(a, f, r, q, n_cells, packed, signals); break;
        case kStorageBF16:
            k_apply_source_records<kStorageBF16><<<g, 64, 0, s>>>(a, f, r, q, n_cells, packed, signals);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace emfd
