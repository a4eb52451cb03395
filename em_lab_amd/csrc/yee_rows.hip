// Fused leapfrog, row-wavefront form: B^{n+1/2} and E^{n+1} in one pass without block barriers.
//
// One wavefront owns one x-row y of one z-chunk and marches up the chunk. Lane l holds VPL
// consecutive cells x = l*VPL .. l*VPL+VPL-1 of the row (float4 / float2 loads when the row is
// full), so the whole row sits in the wave's registers and the x-neighbours it needs come from
// the next / previous lane through DPP wave shifts: no LDS, no __syncthreads, no x-halo (the row
// ends are the domain faces). Per plane z the wave reads
//   E(y, z+1)          the next plane of its own row (kept for the next iteration)
//   E(y+1, z) x,z      the row above, for dEz/dy and dEx/dy of B(y)
//   E(y-1, z), E(y-1, z+1) y, B(y-1, z) x,z   to recompute B^{n+1/2}(y-1) (x and z components),
//                      the -y neighbour the E update needs, instead of exchanging it
//   B(y, z)
// and carries B^{n+1/2}(y, z-1) as the -z neighbour. The neighbour rows are the rows of the other
// waves of the same XCD at the same plane, so they come from L2; HBM sees E and B read once and
// written once. Blocks are ordered so that each XCD owns a contiguous run of (chunk, row) tiles.
//
// Arithmetic is that of faraday() / edge_coefficients() (shaders/update_B.comp:35-69,
// shaders/update_E.comp:31-66), so this kernel is bit-identical to the two-pass and LDS forms.
#include "yee_device.h"

#include <algorithm>

namespace emfd {
namespace {

constexpr int kXcds = 8;
constexpr int kWavesPerBlock = 4;

// lane i receives lane i+1's value; lane 63 receives 0 (DPP wave_shl:1, bound_ctrl)
__device__ __forceinline__ float from_next_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
// lane i receives lane i-1's value; lane 0 receives 0 (DPP wave_shr:1, bound_ctrl)
__device__ __forceinline__ float from_prev_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

template <int VPL, bool FULL>
__device__ __forceinline__ void load_row(const float *__restrict__ p, int xb, int nx, float (&o)[VPL]) {
    if constexpr (FULL && VPL == 4) {
        const float4 v = *reinterpret_cast<const float4 *>(p + xb);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else if constexpr (FULL && VPL == 2) {
        const float2 v = *reinterpret_cast<const float2 *>(p + xb);
        o[0] = v.x; o[1] = v.y;
    } else {
#pragma unroll
        for (int j = 0; j < VPL; ++j) o[j] = xb + j < nx ? p[xb + j] : 0.0f;
    }
}

template <int VPL, bool FULL>
__device__ __forceinline__ void store_row(float *__restrict__ p, int xb, int nx, const float (&v)[VPL]) {
    if constexpr (FULL && VPL == 4) {
        *reinterpret_cast<float4 *>(p + xb) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (FULL && VPL == 2) {
        *reinterpret_cast<float2 *>(p + xb) = make_float2(v[0], v[1]);
    } else {
#pragma unroll
        for (int j = 0; j < VPL; ++j)
            if (xb + j < nx) p[xb + j] = v[j];
    }
}

template <int VPL, bool FULL>
__device__ __forceinline__ void load_ids(const LatticeArgs &a, size_t row, int xb, int nx, uint32_t (&o)[VPL]) {
    if (!a.mat_id) {
#pragma unroll
        for (int j = 0; j < VPL; ++j) o[j] = 0u;
        return;
    }
    if constexpr (FULL && VPL == 4) {
        if (a.id_bytes == 2) {
            const uint2 v = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(a.mat_id) + row + xb);
            o[0] = v.x & 0xffffu; o[1] = v.x >> 16; o[2] = v.y & 0xffffu; o[3] = v.y >> 16;
            return;
        }
    }
#pragma unroll
    for (int j = 0; j < VPL; ++j) o[j] = xb + j < nx ? load_id(a, row + xb + j) : 0u;
}

// Everything one plane iteration reads, loaded one iteration ahead (software pipelining: the
// next plane's loads are in flight while the current plane is computed).
template <int VPL>
struct PlaneIn {
    float eN[3][VPL];  // E(y, z+1)
    float eU[2][VPL];  // E(y+1, z): x, z components
    float eD[3][VPL];  // E(y-1, z)
    float eDn[VPL];    // E(y-1, z+1): y component
    float bD[2][VPL];  // B(y-1, z): x, z components
    float bO[3][VPL];  // B(y, z)
    uint32_t id[VPL], idD[VPL];
};

template <int VPL, bool FULL>
__device__ __forceinline__ void load_plane(const LatticeArgs &a, const FieldPtrs &src, int y, int z, int xb,
                                           size_t sy, size_t sz, bool has_up, bool has_dn, PlaneIn<VPL> &L) {
    const int nx = a.nx, nz = a.nz;
    const size_t row = z * sz + y * sy;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (z + 1 < nz) load_row<VPL, FULL>(src.e[c] + row + sz, xb, nx, L.eN[c]);
        else
#pragma unroll
            for (int j = 0; j < VPL; ++j) L.eN[c][j] = 0.0f;
        load_row<VPL, FULL>(src.b[c] + row, xb, nx, L.bO[c]);
    }
    if (has_up) {
        load_row<VPL, FULL>(src.e[0] + row + sy, xb, nx, L.eU[0]);
        load_row<VPL, FULL>(src.e[2] + row + sy, xb, nx, L.eU[1]);
    } else {
#pragma unroll
        for (int j = 0; j < VPL; ++j) L.eU[0][j] = L.eU[1][j] = 0.0f;
    }
    if (has_dn) {
#pragma unroll
        for (int c = 0; c < 3; ++c) load_row<VPL, FULL>(src.e[c] + row - sy, xb, nx, L.eD[c]);
        if (z + 1 < nz) load_row<VPL, FULL>(src.e[1] + row - sy + sz, xb, nx, L.eDn);
        else
#pragma unroll
            for (int j = 0; j < VPL; ++j) L.eDn[j] = 0.0f;
        load_row<VPL, FULL>(src.b[0] + row - sy, xb, nx, L.bD[0]);
        load_row<VPL, FULL>(src.b[2] + row - sy, xb, nx, L.bD[1]);
        load_ids<VPL, FULL>(a, row - sy, xb, nx, L.idD);
    } else {
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            L.eD[0][j] = L.eD[1][j] = L.eD[2][j] = L.eDn[j] = L.bD[0][j] = L.bD[1][j] = 0.0f;
            L.idD[j] = 0u;
        }
    }
    load_ids<VPL, FULL>(a, row, xb, nx, L.id);
}

template <int VPL, bool FULL, bool PIPE>
__global__ void __launch_bounds__(64 * kWavesPerBlock) k_yee_rows(LatticeArgs a, FieldPtrs src, FieldPtrs dst,
                                                                    int zchunk, int n_tiles) {
    const int lane = threadIdx.x & 63;
    // the wave index is uniform across the wave: say so, so that the tile, row and plane indices
    // (and every branch and address base derived from them) live in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int per = static_cast<int>(gridDim.x) / kXcds;  // gridDim.x is a multiple of kXcds
    const int blk = (static_cast<int>(blockIdx.x) % kXcds) * per + static_cast<int>(blockIdx.x) / kXcds;
    const int tile = blk * kWavesPerBlock + wave;
    if (tile >= n_tiles) return;  // wave-uniform; the kernel has no block-level barrier
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int y = tile % ny;
    const int z0 = (tile / ny) * zchunk;
    const int z1 = min(z0 + zchunk, nz);
    const int zs = z0 > 0 ? z0 - 1 : 0;
    const size_t sy = static_cast<size_t>(nx), sz = static_cast<size_t>(nx) * ny;
    const int xb = lane * VPL;
    const bool has_up = y + 1 < ny, has_dn = y > 0;

    float eC[3][VPL];     // E^n(y, z)
    float bprev[2][VPL];  // B^{n+1/2}(y, z-1): x, y components
#pragma unroll
    for (int c = 0; c < 3; ++c) load_row<VPL, FULL>(src.e[c] + zs * sz + y * sy, xb, nx, eC[c]);
#pragma unroll
    for (int j = 0; j < VPL; ++j) bprev[0][j] = bprev[1][j] = 0.0f;

    PlaneIn<VPL> cur;
    if constexpr (PIPE) load_plane<VPL, FULL>(a, src, y, zs, xb, sy, sz, has_up, has_dn, cur);
    for (int z = zs; z < z1; ++z) {
        const size_t row = z * sz + y * sy;
        PlaneIn<VPL> nxt;
        if constexpr (PIPE) {
            if (z + 1 < z1) load_plane<VPL, FULL>(a, src, y, z + 1, xb, sy, sz, has_up, has_dn, nxt);
        } else {
            load_plane<VPL, FULL>(a, src, y, z, xb, sy, sz, has_up, has_dn, cur);
        }

        // ---- x+1 neighbours: inside the lane, or element 0 of the next lane
        const float nxt_ey = from_next_lane(eC[1][0]), nxt_ez = from_next_lane(eC[2][0]);
        const float nxt_dy = from_next_lane(cur.eD[1][0]);

        // ---- B^{n+1/2}(y, z)
        float bN[3][VPL];
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const float e0[3] = {eC[0][j], eC[1][j], eC[2][j]};
            const float exn[3] = {0.0f, j + 1 < VPL ? eC[1][j + 1] : nxt_ey, j + 1 < VPL ? eC[2][j + 1] : nxt_ez};
            const float eyn[3] = {cur.eU[0][j], 0.0f, cur.eU[1][j]};
            const float ezn[3] = {cur.eN[0][j], cur.eN[1][j], 0.0f};
            float b[3] = {cur.bO[0][j], cur.bO[1][j], cur.bO[2][j]};
            if (xb + j < nx) faraday(a, xb + j, y, z, e0, exn, eyn, ezn, cur.id[j], b);
            else b[0] = b[1] = b[2] = 0.0f;
            bN[0][j] = b[0];
            bN[1][j] = b[1];
            bN[2][j] = b[2];
        }
        // ---- B^{n+1/2}(y-1, z): x and z components (the -y neighbour of E(y)); 0 below y = 0
        float bDN[2][VPL];
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const float e0[3] = {cur.eD[0][j], cur.eD[1][j], cur.eD[2][j]};
            const float exn[3] = {0.0f, j + 1 < VPL ? cur.eD[1][j + 1] : nxt_dy, 0.0f};
            const float eyn[3] = {eC[0][j], 0.0f, eC[2][j]};
            const float ezn[3] = {0.0f, cur.eDn[j], 0.0f};
            float b[3] = {cur.bD[0][j], 0.0f, cur.bD[1][j]};
            if (has_dn && xb + j < nx) faraday(a, xb + j, y - 1, z, e0, exn, eyn, ezn, cur.idD[j], b);
            else b[0] = b[2] = 0.0f;
            bDN[0][j] = b[0];
            bDN[1][j] = b[2];
        }
        // ---- E^{n+1}(y, z)
        if (z >= z0) {
            const float prv_by = from_prev_lane(bN[1][VPL - 1]), prv_bz = from_prev_lane(bN[2][VPL - 1]);
            float eO[3][VPL];
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                const float bxm_y = j > 0 ? bN[1][j - 1] : prv_by;
                const float bxm_z = j > 0 ? bN[2][j - 1] : prv_bz;
                const float bym_x = bDN[0][j], bym_z = bDN[1][j];
                const float bzm_x = bprev[0][j], bzm_y = bprev[1][j];
                const float b0x = bN[0][j], b0y = bN[1][j], b0z = bN[2][j];
                const float curl[3] = {(b0z - bym_z) - (b0y - bzm_y), (b0x - bzm_x) - (b0z - bxm_z),
                                       (b0y - bxm_y) - (b0x - bym_x)};
                float ca[3], cb[3];
                edge_coefficients(a, min(xb + j, nx - 1), y, z, cur.id[j], ca, cb);
#pragma unroll
                for (int c = 0; c < 3; ++c) eO[c][j] = ca[c] * eC[c][j] + cb[c] * curl[c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                store_row<VPL, FULL>(dst.e[c] + row, xb, nx, eO[c]);
                store_row<VPL, FULL>(dst.b[c] + row, xb, nx, bN[c]);
            }
        }
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            bprev[0][j] = bN[0][j];
            bprev[1][j] = bN[1][j];
#pragma unroll
            for (int c = 0; c < 3; ++c) eC[c][j] = cur.eN[c][j];
        }
        if constexpr (PIPE) cur = nxt;
    }
}

template <int VPL, bool FULL>
hipError_t launch_rows_cfg(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int zchunk, bool pipe,
                           hipStream_t s) {
    const int nchunks = (a.nz + zchunk - 1) / zchunk;
    const int n_tiles = a.ny * nchunks;
    int blocks = (n_tiles + kWavesPerBlock - 1) / kWavesPerBlock;
    blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    if (pipe)
        k_yee_rows<VPL, FULL, true><<<blocks, 64 * kWavesPerBlock, 0, s>>>(a, src, dst, zchunk, n_tiles);
    else
        k_yee_rows<VPL, FULL, false><<<blocks, 64 * kWavesPerBlock, 0, s>>>(a, src, dst, zchunk, n_tiles);
    return hipGetLastError();
}

}  // namespace

int yee_rows_supported(const LatticeArgs &a) { return a.nx <= 256 ? 1 : 0; }

// param: bits 0..15 z-chunk (0 = automatic: about 2048 row-waves in flight, 8 per CU over 256 CUs,
// at least 8 planes), bit 16 = software-pipelined plane loads
hipError_t launch_yee_rows(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int param,
                           hipStream_t s) {
    if (!yee_rows_supported(a)) return hipErrorInvalidValue;
    int zchunk = param & 0xffff;
    const bool pipe = (param >> 16) & 1;
    if (zchunk <= 0) {
        int nch = (2048 + a.ny - 1) / a.ny;
        nch = std::max(1, std::min(nch, a.nz / 8));
        zchunk = (a.nz + nch - 1) / nch;
    }
    const int vpl = (a.nx + 63) / 64;
    switch (vpl) {
    case 1:
        return a.nx == 64 ? launch_rows_cfg<1, true>(a, src, dst, zchunk, pipe, s) : launch_rows_cfg<1, false>(a, src, dst, zchunk, pipe, s);
    case 2:
        return a.nx == 128 ? launch_rows_cfg<2, true>(a, src, dst, zchunk, pipe, s) : launch_rows_cfg<2, false>(a, src, dst, zchunk, pipe, s);
    case 3:
        return launch_rows_cfg<3, false>(a, src, dst, zchunk, pipe, s);
    default:
        return a.nx == 256 ? launch_rows_cfg<4, true>(a, src, dst, zchunk, pipe, s) : launch_rows_cfg<4, false>(a, src, dst, zchunk, pipe, s);
    }
}

}  // namespace emfd
