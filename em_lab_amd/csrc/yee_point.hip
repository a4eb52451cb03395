// Fused leapfrog, point form: one thread per texel computes E^{n+1}(p) in a single pass by
// rebuilding the B^{n+1/2} components it needs at p, p-x, p-y and p-z from E^n and B^{n-1/2}
// (ping-pong buffers, so no texel is read after it is overwritten). Full occupancy and no
// synchronisation; the neighbour reads are served by L1/L2, HBM sees each E and B value read once
// and written once. Bit-identical to the other forms (same faraday / coefficient arithmetic).
#include "yee_device.h"

namespace emfd {
namespace {

struct View {
    const LatticeArgs &a;
    const FieldPtrs &f;
    size_t sy, sz;
    __device__ __forceinline__ float e(int c, int x, int y, int z) const {
        // E beyond the upper faces is 0 (shaders/update_B.comp:26-29)
        return (x < a.nx && y < a.ny && z < a.nz) ? f.e[c][z * sz + y * sy + x] : 0.0f;
    }
    __device__ __forceinline__ float b(int c, int x, int y, int z) const { return f.b[c][z * sz + y * sy + x]; }
    __device__ __forceinline__ uint32_t id(int x, int y, int z) const {
        return a.mat_id ? load_id(a, z * sz + y * sy + x) : 0u;
    }
};

__global__ void __launch_bounds__(256) k_yee_point(LatticeArgs a, FieldPtrs src, FieldPtrs dst) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= a.nx || y >= a.ny) return;
    const View v{a, src, static_cast<size_t>(a.nx), static_cast<size_t>(a.nx) * a.ny};
    const size_t p = z * v.sz + y * v.sy + x;
    const float ex = v.e(0, x, y, z), ey = v.e(1, x, y, z), ez = v.e(2, x, y, z);
    const uint32_t id0 = v.id(x, y, z);
    // B^{n+1/2}(p), all components (term0, term1 per update_B.comp:41-46)
    const float exp_y = v.e(1, x + 1, y, z), exp_z = v.e(2, x + 1, y, z);
    const float eyp_x = v.e(0, x, y + 1, z), eyp_z = v.e(2, x, y + 1, z);
    const float ezp_x = v.e(0, x, y, z + 1), ezp_y = v.e(1, x, y, z + 1);
    const float b0x = faraday_component(a, 0, x, y, z, eyp_z - ez, ezp_y - ey, id0, v.b(0, x, y, z));
    const float b0y = faraday_component(a, 1, x, y, z, ezp_x - ex, exp_z - ez, id0, v.b(1, x, y, z));
    const float b0z = faraday_component(a, 2, x, y, z, exp_y - ey, eyp_x - ex, id0, v.b(2, x, y, z));
    // -x neighbour: By, Bz at (x-1, y, z); 0 below the low face
    float bxm_y = 0.0f, bxm_z = 0.0f;
    if (x > 0) {
        const int q = x - 1;
        const float qx = v.e(0, q, y, z), qy = v.e(1, q, y, z), qz = v.e(2, q, y, z);
        const uint32_t idq = v.id(q, y, z);
        bxm_y = faraday_component(a, 1, q, y, z, v.e(0, q, y, z + 1) - qx, ez - qz, idq, v.b(1, q, y, z));
        bxm_z = faraday_component(a, 2, q, y, z, ey - qy, v.e(0, q, y + 1, z) - qx, idq, v.b(2, q, y, z));
    }
    // -y neighbour: Bx, Bz at (x, y-1, z)
    float bym_x = 0.0f, bym_z = 0.0f;
    if (y > 0) {
        const int q = y - 1;
        const float qx = v.e(0, x, q, z), qy = v.e(1, x, q, z), qz = v.e(2, x, q, z);
        const uint32_t idq = v.id(x, q, z);
        bym_x = faraday_component(a, 0, x, q, z, ez - qz, v.e(1, x, q, z + 1) - qy, idq, v.b(0, x, q, z));
        bym_z = faraday_component(a, 2, x, q, z, v.e(1, x + 1, q, z) - qy, ex - qx, idq, v.b(2, x, q, z));
    }
    // -z neighbour: Bx, By at (x, y, z-1)
    float bzm_x = 0.0f, bzm_y = 0.0f;
    if (z > 0) {
        const int q = z - 1;
        const float qx = v.e(0, x, y, q), qy = v.e(1, x, y, q), qz = v.e(2, x, y, q);
        const uint32_t idq = v.id(x, y, q);
        bzm_x = faraday_component(a, 0, x, y, q, v.e(2, x, y + 1, q) - qz, ey - qy, idq, v.b(0, x, y, q));
        bzm_y = faraday_component(a, 1, x, y, q, ex - qx, v.e(2, x + 1, y, q) - qz, idq, v.b(1, x, y, q));
    }
    const float curl[3] = {(b0z - bym_z) - (b0y - bzm_y), (b0x - bzm_x) - (b0z - bxm_z), (b0y - bxm_y) - (b0x - bym_x)};
    float ca[3], cb[3];
    edge_coefficients(a, x, y, z, id0, ca, cb);
    dst.e[0][p] = ca[0] * ex + cb[0] * curl[0];
    dst.e[1][p] = ca[1] * ey + cb[1] * curl[1];
    dst.e[2][p] = ca[2] * ez + cb[2] * curl[2];
    dst.b[0][p] = b0x;
    dst.b[1][p] = b0y;
    dst.b[2][p] = b0z;
}

}  // namespace

hipError_t launch_yee_point(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, hipStream_t s) {
    dim3 block(64, 4, 1), grid((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
    k_yee_point<<<grid, block, 0, s>>>(a, src, dst);
    return hipGetLastError();
}

}  // namespace emfd
