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

constexpr int kXcds = 8;  // MI355X: 8 accelerator dies, each with its own L2

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

// ---------------------------------------------------------------------------------------------
// Fused leapfrog: B^{n+1/2} and E^{n+1} in one pass, ping-pong buffers.
//
// A workgroup owns a TX x (TY*RY) tile of (x, y) columns and marches up a z-chunk. Per plane z it
//   - holds E^n(z) of the tile plus a one-cell halo in LDS (double-buffered: E^n(z+1) is staged
//     while plane z is consumed),
//   - computes B^{n+1/2}(z) on the tile and on the low-x column / low-y row halo (recomputed, not
//     exchanged), keeps it in LDS for the -x/-y neighbours of the E update, and carries the own
//     texel's B(z) in registers as the -z neighbour of plane z+1,
//   - updates E^{n+1}(z) from LDS and registers.
// HBM traffic per texel is E and B read once and written once (48 B) plus the class id.
// A z-chunk that does not start at 0 recomputes B at z0-1 first (one extra plane).
// ---------------------------------------------------------------------------------------------
template <int TX, int TY, int RY>
struct FusedTile {
    static constexpr int RYT = TY * RY;            // rows per tile
    static constexpr int EW = TX + 2, EH = RYT + 2;  // E plane with halo on every side
    static constexpr int BW = TX + 1, BH = RYT + 1;  // B plane with low-side halo
    static constexpr int NT = TX * TY;
    static constexpr int EHALO = 2 * TX + 2 * EH;  // top/bottom rows + left/right columns
    static constexpr int BHALO = TX + RYT;         // low-y row + low-x column
    static constexpr int EH_SLOTS = (EHALO + NT - 1) / NT;
    static constexpr int BH_SLOTS = (BHALO + NT - 1) / NT;
};

template <int TX, int TY, int RY>
__global__ void __launch_bounds__(TX *TY) k_yee_fused(LatticeArgs a, FieldPtrs src, FieldPtrs dst, int tiles_x,
                                                       int tiles_y, int zchunk, int n_tiles) {
    using T = FusedTile<TX, TY, RY>;
    constexpr int RYT = T::RYT, EW = T::EW, EH = T::EH, BW = T::BW, NT = T::NT;
    __shared__ float sE[2][3][EH][EW];
    __shared__ float sB[3][T::BH][BW];

    // XCD-aware tile order: consecutive blocks land on different XCDs, so give each XCD a
    // contiguous run of tiles (neighbours then share that XCD's L2 for the halo reads).
    const int per = (n_tiles + kXcds - 1) / kXcds;
    const int tile = (static_cast<int>(blockIdx.x) % kXcds) * per + static_cast<int>(blockIdx.x) / kXcds;
    if (tile >= n_tiles) return;
    const int ti_x = tile % tiles_x;
    const int ti_y = (tile / tiles_x) % tiles_y;
    const int ti_z = tile / (tiles_x * tiles_y);
    const int x0 = ti_x * TX, y0 = ti_y * RYT;
    const int z0 = ti_z * zchunk;
    const int z1 = min(z0 + zchunk, a.nz);
    const int zs = z0 > 0 ? z0 - 1 : 0;

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = tx + TX * ty;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t sy = static_cast<size_t>(nx), sz = static_cast<size_t>(nx) * ny;

    // ---- E-region entry coordinates (own rows + halo) ----
    // own: (ex, ey) = (tx+1, ty + TY r + 1); halo entries enumerated by h.
    int ehx[T::EH_SLOTS], ehy[T::EH_SLOTS];
#pragma unroll
    for (int s = 0; s < T::EH_SLOTS; ++s) {
        const int h = tid + s * NT;
        int ex = -1, ey = -1;
        if (h < TX) { ex = h + 1; ey = 0; }
        else if (h < 2 * TX) { ex = h - TX + 1; ey = EH - 1; }
        else if (h < T::EHALO) {
            const int hh = h - 2 * TX;
            ex = hh < EH ? 0 : EW - 1;
            ey = hh % EH;
        }
        ehx[s] = ex;
        ehy[s] = ey;
    }
    // ---- B halo entries: low-y row (bx in 1..TX, by 0), low-x column (bx 0, by 1..RYT) ----
    int bhx[T::BH_SLOTS], bhy[T::BH_SLOTS];
#pragma unroll
    for (int s = 0; s < T::BH_SLOTS; ++s) {
        const int h = tid + s * NT;
        int bx = -1, by = -1;
        if (h < TX) { bx = h + 1; by = 0; }
        else if (h < T::BHALO) { bx = 0; by = h - TX + 1; }
        bhx[s] = bx;
        bhy[s] = by;
    }

    auto ld_e = [&](int c, int gx, int gy, int gz) -> float {
        return (gx >= 0 && gx < nx && gy >= 0 && gy < ny && gz < nz) ? src.e[c][gz * sz + gy * sy + gx] : 0.0f;
    };
    auto ld_b = [&](int c, int gx, int gy, int gz) -> float {
        return (gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? src.b[c][gz * sz + gy * sy + gx] : 0.0f;
    };
    auto ld_id = [&](int gx, int gy, int gz) -> uint32_t {
        return (a.mat_id && gx >= 0 && gx < nx && gy >= 0 && gy < ny) ? load_id(a, gz * sz + gy * sy + gx) : 0u;
    };

    // registers: own E at z (current) and staged E at z+1, own B_old(z), halo E staged
    float eo[RY][3], en[RY][3];
    float bprev[RY][2];  // B^{n+1/2} at (own, z-1): x, y components (the -z neighbour)

    // prologue: E plane zs into LDS buffer (zs & 1) and registers
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int gy = y0 + ty + TY * r;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            eo[r][c] = ld_e(c, x0 + tx, gy, zs);
            sE[zs & 1][c][ty + TY * r + 1][tx + 1] = eo[r][c];
        }
        bprev[r][0] = bprev[r][1] = 0.0f;
    }
#pragma unroll
    for (int s = 0; s < T::EH_SLOTS; ++s)
        if (ehx[s] >= 0)
#pragma unroll
            for (int c = 0; c < 3; ++c) sE[zs & 1][c][ehy[s]][ehx[s]] = ld_e(c, x0 + ehx[s] - 1, y0 + ehy[s] - 1, zs);

    for (int z = zs; z < z1; ++z) {
        const int cur = z & 1, nxt = cur ^ 1;
        // (a) stage E(z+1) into LDS (zero beyond the top face)
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int gy = y0 + ty + TY * r;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                en[r][c] = ld_e(c, x0 + tx, gy, z + 1);
                sE[nxt][c][ty + TY * r + 1][tx + 1] = en[r][c];
            }
        }
#pragma unroll
        for (int s = 0; s < T::EH_SLOTS; ++s)
            if (ehx[s] >= 0)
#pragma unroll
                for (int c = 0; c < 3; ++c) sE[nxt][c][ehy[s]][ehx[s]] = ld_e(c, x0 + ehx[s] - 1, y0 + ehy[s] - 1, z + 1);
        // B_old(z) for own texels and B-halo texels, class ids
        float bo[RY][3];
        uint32_t ido[RY];
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int gy = y0 + ty + TY * r;
#pragma unroll
            for (int c = 0; c < 3; ++c) bo[r][c] = ld_b(c, x0 + tx, gy, z);
            ido[r] = ld_id(x0 + tx, gy, z);
        }
        float bho[T::BH_SLOTS][3];
        uint32_t idh[T::BH_SLOTS];
#pragma unroll
        for (int s = 0; s < T::BH_SLOTS; ++s) {
            const int gx = x0 + bhx[s] - 1, gy = y0 + bhy[s] - 1;
#pragma unroll
            for (int c = 0; c < 3; ++c) bho[s][c] = bhx[s] >= 0 ? ld_b(c, gx, gy, z) : 0.0f;
            idh[s] = bhx[s] >= 0 ? ld_id(gx, gy, z) : 0u;
        }
        __syncthreads();  // #1: E(z+1) staged

        // (d) B^{n+1/2}(z): own texels
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const int ey = ty + TY * r + 1, ex = tx + 1;
            const int gx = x0 + tx, gy = y0 + ty + TY * r;
            float b[3] = {bo[r][0], bo[r][1], bo[r][2]};
            if (gx < nx && gy < ny) {
                float e0[3], exn[3], eyn[3], ezn[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    e0[c] = sE[cur][c][ey][ex];
                    exn[c] = sE[cur][c][ey][ex + 1];
                    eyn[c] = sE[cur][c][ey + 1][ex];
                    ezn[c] = en[r][c];
                }
                faraday(a, gx, gy, z, e0, exn, eyn, ezn, ido[r], b);
                if (z >= z0) {
                    const size_t p = z * sz + gy * sy + gx;
#pragma unroll
                    for (int c = 0; c < 3; ++c) dst.b[c][p] = b[c];
                }
            } else {
                b[0] = b[1] = b[2] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) sB[c][ey][ex] = b[c];
            bo[r][0] = b[0];
            bo[r][1] = b[1];
            bo[r][2] = b[2];
        }
        // halo texels (recomputed; zero outside the domain = "B below the low faces is 0")
#pragma unroll
        for (int s = 0; s < T::BH_SLOTS; ++s) {
            if (bhx[s] < 0) continue;
            const int bx = bhx[s], by = bhy[s];
            const int gx = x0 + bx - 1, gy = y0 + by - 1;
            float b[3] = {bho[s][0], bho[s][1], bho[s][2]};
            if (gx >= 0 && gy >= 0 && gx < nx && gy < ny) {
                float e0[3], exn[3], eyn[3], ezn[3];
                // E index of texel (gx, gy) is (bx, by) in the E region
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    e0[c] = sE[cur][c][by][bx];
                    exn[c] = sE[cur][c][by][bx + 1];
                    eyn[c] = sE[cur][c][by + 1][bx];
                    ezn[c] = sE[nxt][c][by][bx];
                }
                faraday(a, gx, gy, z, e0, exn, eyn, ezn, idh[s], b);
            } else {
                b[0] = b[1] = b[2] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) sB[c][by][bx] = b[c];
        }
        __syncthreads();  // #2: B(z) visible

        // (f) E^{n+1}(z) for own texels
        if (z >= z0) {
#pragma unroll
            for (int r = 0; r < RY; ++r) {
                const int gx = x0 + tx, gy = y0 + ty + TY * r;
                if (gx < nx && gy < ny) {
                    const int by = ty + TY * r + 1, bx = tx + 1;
                    const float b0x = bo[r][0], b0y = bo[r][1], b0z = bo[r][2];
                    const float bxm_y = sB[1][by][bx - 1], bxm_z = sB[2][by][bx - 1];
                    const float bym_x = sB[0][by - 1][bx], bym_z = sB[2][by - 1][bx];
                    const float bzm_x = bprev[r][0], bzm_y = bprev[r][1];
                    const float curl[3] = {(b0z - bym_z) - (b0y - bzm_y), (b0x - bzm_x) - (b0z - bxm_z),
                                           (b0y - bxm_y) - (b0x - bym_x)};
                    float ca[3], cb[3];
                    edge_coefficients(a, gx, gy, z, ido[r], ca, cb);
                    const size_t p = z * sz + gy * sy + gx;
#pragma unroll
                    for (int c = 0; c < 3; ++c) dst.e[c][p] = ca[c] * eo[r][c] + cb[c] * curl[c];
                }
            }
        }
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            bprev[r][0] = bo[r][0];
            bprev[r][1] = bo[r][1];
#pragma unroll
            for (int c = 0; c < 3; ++c) eo[r][c] = en[r][c];
        }
        // next iteration writes sE[cur] (read in (d) above, before barrier #2) and sB (read in
        // (f) above, before the next barrier #1): no third barrier needed.
    }
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

hipError_t launch_yee_two_pass(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s) {
    dim3 block(64, 4, 1), grid((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
    k_update_b<<<grid, block, 0, s>>>(a, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_update_e<<<grid, block, 0, s>>>(a, f);
    return hipGetLastError();
}

namespace {
template <int TX, int TY, int RY>
hipError_t launch_fused_cfg(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int zchunk,
                            hipStream_t s) {
    const int tiles_x = (a.nx + TX - 1) / TX;
    const int tiles_y = (a.ny + TY * RY - 1) / (TY * RY);
    const int tiles_z = (a.nz + zchunk - 1) / zchunk;
    const int n_tiles = tiles_x * tiles_y * tiles_z;
    const int blocks = (n_tiles + kXcds - 1) / kXcds * kXcds;
    dim3 block(TX, TY, 1);
    k_yee_fused<TX, TY, RY><<<blocks, block, 0, s>>>(a, src, dst, tiles_x, tiles_y, zchunk, n_tiles);
    return hipGetLastError();
}
}  // namespace

// tile_cfg: low byte = tile shape id, upper bits = z-chunk (0 = choose automatically)
hipError_t launch_yee_fused(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int tile_cfg,
                            hipStream_t s) {
    const int shape = tile_cfg & 0xff;
    int zchunk = tile_cfg >> 8;
    auto auto_chunk = [&](int tx, int ty) {
        // aim for >= 4 workgroups per CU over 256 CUs
        const long cols = static_cast<long>((a.nx + tx - 1) / tx) * ((a.ny + ty - 1) / ty);
        long chunks = (2048 + cols - 1) / cols;
        if (chunks < 1) chunks = 1;
        int zc = static_cast<int>((a.nz + chunks - 1) / chunks);
        return std::max(zc, 8);
    };
    switch (shape) {
    case 1:
        if (!zchunk) zchunk = auto_chunk(64, 8);
        return launch_fused_cfg<64, 4, 2>(a, src, dst, zchunk, s);
    case 2:
        if (!zchunk) zchunk = auto_chunk(64, 16);
        return launch_fused_cfg<64, 4, 4>(a, src, dst, zchunk, s);
    case 3:
        if (!zchunk) zchunk = auto_chunk(64, 8);
        return launch_fused_cfg<64, 8, 1>(a, src, dst, zchunk, s);
    case 4:
        if (!zchunk) zchunk = auto_chunk(128, 4);
        return launch_fused_cfg<128, 2, 2>(a, src, dst, zchunk, s);
    default:
        if (!zchunk) zchunk = auto_chunk(64, 8);
        return launch_fused_cfg<64, 4, 2>(a, src, dst, zchunk, s);
    }
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
