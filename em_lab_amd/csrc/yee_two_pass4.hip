// Two-pass leapfrog, vectorised: update_B then update_E in place (the reference's dispatch order,
// src/SimulationGrid.cpp:393-408), each thread owning four consecutive x cells (float4 streams,
// 1 KB per wavefront instruction). Requires nx % 4 == 0.
#include "yee_quad.h"

namespace emfd {
namespace {

// Block (64, 4): wavefront w = row y0 + w, 256 cells; its absorber coefficients are staged in
// its own LDS region (coef_rows.h).
template <int ST>
__global__ void __launch_bounds__(256) k_update_b4(LatticeArgs a, FieldPtrs f, int z_base) {
    extern __shared__ float coef_lds[];
    const int x0 = (blockIdx.x * 64 + threadIdx.x) * 4, y = blockIdx.y * 4 + threadIdx.y, z = z_base + blockIdx.z;
    float *rows = coef_lds + threadIdx.y * coef_entries(a, 1) * kCoefStride;
    b_quad_rows<ST>(a, f, x0, y, z, x0 < a.nx && y < a.ny, rows, threadIdx.x);
}

// PM: see edge_coefficients (false only for lattices without a material texel in the absorber)
template <int ST, bool PM>
__global__ void __launch_bounds__(256) k_update_e4(LatticeArgs a, FieldPtrs f, int z_base) {
    extern __shared__ float coef_lds[];
    const int x0 = (blockIdx.x * 64 + threadIdx.x) * 4, y = blockIdx.y * 4 + threadIdx.y, z = z_base + blockIdx.z;
    float *rows = coef_lds + threadIdx.y * coef_entries(a, 1) * kCoefStride;
    e_quad_rows<ST, PM>(a, f, x0, y, z, x0 < a.nx && y < a.ny, rows, threadIdx.x);
}

template <int ST>
hipError_t launch_st(const LatticeArgs &a, const FieldPtrs &f, int z0, int nzs, hipStream_t s, int passes) {
    dim3 block(64, 4, 1), grid((a.nx / 4 + 63) / 64, (a.ny + 3) / 4, nzs);
    const size_t lds = static_cast<size_t>(4 * coef_entries(a, 1) * kCoefStride) * sizeof(float);
    if (passes & 1) {
        k_update_b4<ST><<<grid, block, lds, s>>>(a, f, z0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (!(passes & 2)) return hipSuccess;
    if (a.pml_materials)
        k_update_e4<ST, true><<<grid, block, lds, s>>>(a, f, z0);
    else
        k_update_e4<ST, false><<<grid, block, lds, s>>>(a, f, z0);
    return hipGetLastError();
}

}  // namespace

int yee_two_pass4_supported(const LatticeArgs &a) { return a.nx % 4 == 0 ? 1 : 0; }

// Non-temporal hints on the stores and on the loads of the plane each pass reads once (ST = 2 of
// yee_quad.h's streaming levels, the fastest measured).
hipError_t launch_yee_two_pass4(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s, int passes) {
    if (!yee_two_pass4_supported(a)) return hipErrorInvalidValue;
    return launch_st<2>(a, f, 0, a.nz, s, passes);
}

}  // namespace emfd
