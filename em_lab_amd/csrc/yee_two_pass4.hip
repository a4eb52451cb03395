// Two-pass leapfrog, vectorised: update_B then update_E in place (the reference's dispatch order,
// src/SimulationGrid.cpp:393-408), each thread owning four consecutive x cells (float4 streams,
// 1 KB per wavefront instruction). Requires nx % 4 == 0.
#include "yee_quad.h"

namespace emfd {
namespace {

__global__ void __launch_bounds__(256) k_update_b4(LatticeArgs a, FieldPtrs f) {
    const int x0 = (blockIdx.x * 64 + threadIdx.x) * 4, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x0 >= a.nx || y >= a.ny) return;
    b_quad(a, f, x0, y, z);
}

__global__ void __launch_bounds__(256) k_update_e4(LatticeArgs a, FieldPtrs f) {
    const int x0 = (blockIdx.x * 64 + threadIdx.x) * 4, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x0 >= a.nx || y >= a.ny) return;
    e_quad(a, f, x0, y, z);
}

}  // namespace

int yee_two_pass4_supported(const LatticeArgs &a) { return a.nx % 4 == 0 ? 1 : 0; }

hipError_t launch_yee_two_pass4(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s) {
    if (!yee_two_pass4_supported(a)) return hipErrorInvalidValue;
    dim3 block(64, 4, 1), grid((a.nx / 4 + 63) / 64, (a.ny + 3) / 4, a.nz);
    k_update_b4<<<grid, block, 0, s>>>(a, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_update_e4<<<grid, block, 0, s>>>(a, f);
    return hipGetLastError();
}

}  // namespace emfd
