// Fused leapfrog, z-march with row groups: B^{n+1/2} and E^{n+1} in one pass over the planes of a
// z-chunk, so that HBM sees E and B read once and written once (ping-pong buffers).
//
// A workgroup owns R consecutive rows y0 .. y0+R-1 of one z-chunk and one helper wavefront.
// Compute wave w holds row y0+w whole (nx <= 256: lane l holds cells 4l .. 4l+3), so the x
// neighbours come from the next / previous lane through DPP shifts. The y neighbours are the
// other waves' rows, exchanged through LDS: per plane z each wave publishes its E(z) x/z
// components (B^{n+1/2}(y) needs E(y+1)), then its B^{n+1/2}(z) x/z components (E^{n+1}(y) needs
// B(y-1)). The helper wave supplies the two rows the group does not own: E(y0+R, z) above it and
// B^{n+1/2}(y0-1, z) below it (rebuilt from E and B rows that neighbouring groups read anyway,
// so they come from L2). B^{n+1/2}(y, z-1), the -z neighbour, stays in registers; the next
// plane's E is loaded one iteration ahead. A chunk starts one plane early to rebuild
// B^{n+1/2}(z0-1) without storing it.
//
// Per-texel arithmetic is faraday() / edge_coefficients(), so the result is bit-identical to the
// two-pass kernels.
#include "yee_device.h"

namespace emfd {
namespace {

constexpr int kRows = 4;         // compute waves (rows) per workgroup
constexpr int kXcdsZ = 8;

__device__ __forceinline__ float zm_next_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float zm_prev_lane(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

__device__ __forceinline__ void zm_load4(const float *p, int xb, int nx, float (&o)[4]) {
    if (xb + 4 <= nx) {
        const float4 v = *reinterpret_cast<const float4 *>(p + xb);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else {
        o[0] = o[1] = o[2] = o[3] = 0.0f;
    }
}
__device__ __forceinline__ void zm_store4(float *p, int xb, int nx, const float (&v)[4]) {
    if (xb + 4 <= nx) *reinterpret_cast<float4 *>(p + xb) = make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ void zm_ids4(const LatticeArgs &a, size_t row, int xb, int nx, uint32_t (&o)[4]) {
    if (!a.mat_id || xb + 4 > nx) {
        o[0] = o[1] = o[2] = o[3] = 0u;
    } else if (a.id_bytes == 2) {
        const uint2 v = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(a.mat_id) + row + xb);
        o[0] = v.x & 0xffffu; o[1] = v.x >> 16; o[2] = v.y & 0xffffu; o[3] = v.y >> 16;
    } else {
        const uint4 v = *reinterpret_cast<const uint4 *>(static_cast<const uint32_t *>(a.mat_id) + row + xb);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    }
}

// sE[k]: E(y0 + k, z) x, z components (k = 0 .. R); sB[k]: B^{n+1/2}(y0 - 1 + k, z) x, z (k = 0 .. R)
struct ZmShared {
    float e[kRows + 1][2][256];
    float b[kRows + 1][2][256];
};

__global__ void __launch_bounds__(64 * (kRows + 1)) k_yee_zmarch(LatticeArgs a, FieldPtrs src, FieldPtrs dst,
                                                                 int zchunk, int n_groups, int n_tiles) {
    __shared__ ZmShared sh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int per = static_cast<int>(gridDim.x) / kXcdsZ;
    const int tile = (static_cast<int>(blockIdx.x) % kXcdsZ) * per + static_cast<int>(blockIdx.x) / kXcdsZ;
    if (tile >= n_tiles) return;  // uniform over the workgroup
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int y0 = (tile % n_groups) * kRows;
    const int z0 = (tile / n_groups) * zchunk;
    const int z1 = min(z0 + zchunk, nz);
    const int zs = z0 > 0 ? z0 - 1 : 0;
    const size_t sy = static_cast<size_t>(nx), sz = static_cast<size_t>(nx) * ny;
    const int xb = lane * 4;

    if (wave == kRows) {
        // ---------------- helper wave: E(y0+R) above, B^{n+1/2}(y0-1) below ----------------
        const int yu = y0 + kRows, yd = y0 - 1;
        const bool up = yu < ny, dn = yd >= 0;
        // rows of plane z (cur) and of plane z + 1 (nxt, loaded one iteration ahead)
        struct HelperRows {
            float exU[4], ezU[4];            // E(yu, z) x, z
            float exD[4], ezD[4], eyDn[4];   // E(yd, z) x, z; E(yd, z+1) y
            float bxD[4], bzD[4];            // B(yd, z) x, z
            uint32_t idD[4];
        };
        auto load_rows = [&](int z, HelperRows &r) {
            if (up) {
                zm_load4(src.e[0] + z * sz + yu * sy, xb, nx, r.exU);
                zm_load4(src.e[2] + z * sz + yu * sy, xb, nx, r.ezU);
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) r.exU[j] = r.ezU[j] = 0.0f;
            }
            if (dn) {
                const size_t row = z * sz + yd * sy;
                zm_load4(src.e[0] + row, xb, nx, r.exD);
                zm_load4(src.e[2] + row, xb, nx, r.ezD);
                if (z + 1 < nz) zm_load4(src.e[1] + row + sz, xb, nx, r.eyDn);
                else r.eyDn[0] = r.eyDn[1] = r.eyDn[2] = r.eyDn[3] = 0.0f;
                zm_load4(src.b[0] + row, xb, nx, r.bxD);
                zm_load4(src.b[2] + row, xb, nx, r.bzD);
                zm_ids4(a, row, xb, nx, r.idD);
            }
        };
        float eyD[4];  // Ey(yd, z)
        if (dn) zm_load4(src.e[1] + zs * sz + yd * sy, xb, nx, eyD);
        else eyD[0] = eyD[1] = eyD[2] = eyD[3] = 0.0f;
        HelperRows cur;
        load_rows(zs, cur);
        for (int z = zs; z < z1; ++z) {
            HelperRows nxt;
            if (z + 1 < z1) load_rows(z + 1, nxt);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sh.e[kRows][0][xb + j] = cur.exU[j];
                sh.e[kRows][1][xb + j] = cur.ezU[j];
            }
            __syncthreads();  // (1) every row's E(z) x/z is published
            float bo[2][4] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
            if (dn) {
                const float nxt_ey = zm_next_lane(eyD[0]);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (xb + j >= nx) continue;
                    // B^{n+1/2}(yd): x needs dEz/dy, dEy/dz; z needs dEy/dx, dEx/dy (y unused)
                    const float e0[3] = {cur.exD[j], eyD[j], cur.ezD[j]};
                    const float exn[3] = {0.0f, j < 3 ? eyD[j + 1] : nxt_ey, 0.0f};
                    const float eyn[3] = {sh.e[0][0][xb + j], 0.0f, sh.e[0][1][xb + j]};
                    const float ezn[3] = {0.0f, cur.eyDn[j], 0.0f};
                    float b[3] = {cur.bxD[j], 0.0f, cur.bzD[j]};
                    faraday(a, xb + j, yd, z, e0, exn, eyn, ezn, cur.idD[j], b);
                    bo[0][j] = b[0];
                    bo[1][j] = b[2];
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                sh.b[0][0][xb + j] = bo[0][j];
                sh.b[0][1][xb + j] = bo[1][j];
            }
            __syncthreads();  // (2) every row's B^{n+1/2}(z) x/z is published
#pragma unroll
            for (int j = 0; j < 4; ++j) eyD[j] = cur.eyDn[j];
            if (z + 1 < z1) cur = nxt;
        }
        return;
    }

    // ---------------- compute wave: row y ----------------
    const int y = y0 + wave;
    const bool live = y < ny;
    float eC[3][4], eN[3][4];
    float bp[2][4] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};  // B^{n+1/2}(y, z-1) x, y
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if (live) zm_load4(src.e[c] + zs * sz + y * sy, xb, nx, eC[c]);
        else eC[c][0] = eC[c][1] = eC[c][2] = eC[c][3] = 0.0f;
        if (live && zs + 1 < nz) zm_load4(src.e[c] + (zs + 1) * sz + y * sy, xb, nx, eN[c]);
        else eN[c][0] = eN[c][1] = eN[c][2] = eN[c][3] = 0.0f;
    }
    // B^{n-1/2}(y, z) and the class ids of the plane, one iteration ahead like E
    float bo[3][4];
    uint32_t id[4];
    if (live) {
#pragma unroll
        for (int c = 0; c < 3; ++c) zm_load4(src.b[c] + zs * sz + y * sy, xb, nx, bo[c]);
        zm_ids4(a, zs * sz + y * sy, xb, nx, id);
    }
    for (int z = zs; z < z1; ++z) {
        const size_t row = z * sz + y * sy;
        float eNN[3][4], boN[3][4];
        uint32_t idN[4];
        const bool more = live && z + 1 < z1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (more && z + 2 < nz) zm_load4(src.e[c] + row + 2 * sz, xb, nx, eNN[c]);
            else eNN[c][0] = eNN[c][1] = eNN[c][2] = eNN[c][3] = 0.0f;
            if (more) zm_load4(src.b[c] + row + sz, xb, nx, boN[c]);
        }
        if (more) zm_ids4(a, row + sz, xb, nx, idN);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sh.e[wave][0][xb + j] = eC[0][j];
            sh.e[wave][1][xb + j] = eC[2][j];
        }
        __syncthreads();  // (1)
        float bN[3][4];
        {
            const float nxt_ey = zm_next_lane(eC[1][0]), nxt_ez = zm_next_lane(eC[2][0]);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float e0[3] = {eC[0][j], eC[1][j], eC[2][j]};
                const float exn[3] = {0.0f, j < 3 ? eC[1][j + 1] : nxt_ey, j < 3 ? eC[2][j + 1] : nxt_ez};
                const float eyn[3] = {sh.e[wave + 1][0][xb + j], 0.0f, sh.e[wave + 1][1][xb + j]};
                const float ezn[3] = {eN[0][j], eN[1][j], 0.0f};
                float b[3] = {0.0f, 0.0f, 0.0f};
                if (live && xb + j < nx) {
                    b[0] = bo[0][j];
                    b[1] = bo[1][j];
                    b[2] = bo[2][j];
                    faraday(a, xb + j, y, z, e0, exn, eyn, ezn, id[j], b);
                }
                bN[0][j] = b[0];
                bN[1][j] = b[1];
                bN[2][j] = b[2];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sh.b[wave + 1][0][xb + j] = bN[0][j];
            sh.b[wave + 1][1][xb + j] = bN[2][j];
        }
        __syncthreads();  // (2)
        if (live && z >= z0) {
            const float prv_by = zm_prev_lane(bN[1][3]), prv_bz = zm_prev_lane(bN[2][3]);
            float eO[3][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float bxm_y = j > 0 ? bN[1][j - 1] : prv_by;
                const float bxm_z = j > 0 ? bN[2][j - 1] : prv_bz;
                const float bym_x = sh.b[wave][0][xb + j], bym_z = sh.b[wave][1][xb + j];
                const float bzm_x = bp[0][j], bzm_y = bp[1][j];
                const float b0x = bN[0][j], b0y = bN[1][j], b0z = bN[2][j];
                const float curl[3] = {(b0z - bym_z) - (b0y - bzm_y), (b0x - bzm_x) - (b0z - bxm_z),
                                       (b0y - bxm_y) - (b0x - bym_x)};
                float ca[3], cb[3];
                edge_coefficients(a, min(xb + j, nx - 1), y, z, id[j], ca, cb);
#pragma unroll
                for (int c = 0; c < 3; ++c) eO[c][j] = ca[c] * eC[c][j] + cb[c] * curl[c];
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                zm_store4(dst.e[c] + row, xb, nx, eO[c]);
                zm_store4(dst.b[c] + row, xb, nx, bN[c]);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            bp[0][j] = bN[0][j];
            bp[1][j] = bN[1][j];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                eC[c][j] = eN[c][j];
                eN[c][j] = eNN[c][j];
            }
        }
        if (more) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                id[j] = idN[j];
#pragma unroll
                for (int c = 0; c < 3; ++c) bo[c][j] = boN[c][j];
            }
        }
    }
}

}  // namespace

int yee_zmarch_supported(const LatticeArgs &a) { return a.nx <= 256 && a.nx % 4 == 0 ? 1 : 0; }

// zchunk <= 0: chosen so that about three workgroups per CU are in flight (256 CUs)
hipError_t launch_yee_zmarch(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int zchunk,
                             hipStream_t s) {
    if (!yee_zmarch_supported(a)) return hipErrorInvalidValue;
    const int n_groups = (a.ny + kRows - 1) / kRows;
    if (zchunk <= 0) {
        int nch = (768 + n_groups - 1) / n_groups;
        nch = nch < 1 ? 1 : nch;
        if (nch > a.nz / 4) nch = a.nz / 4 > 0 ? a.nz / 4 : 1;
        zchunk = (a.nz + nch - 1) / nch;
    }
    const int nchunks = (a.nz + zchunk - 1) / zchunk;
    const int n_tiles = n_groups * nchunks;
    const int blocks = (n_tiles + kXcdsZ - 1) / kXcdsZ * kXcdsZ;
    k_yee_zmarch<<<blocks, 64 * (kRows + 1), 0, s>>>(a, src, dst, zchunk, n_groups, n_tiles);
    return hipGetLastError();
}

}  // namespace emfd
