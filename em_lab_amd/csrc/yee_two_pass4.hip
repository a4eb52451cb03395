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

// XCD-contiguous order: the hardware deals workgroups round-robin to the 8 XCDs; remapping the
// dispatch index b to tile (b % 8) * per + b / 8 gives each XCD one contiguous run of (y, z)
// blocks, so the +y / +z neighbour rows a block reads were just read by a block of the same XCD
// (its L2) instead of one on another XCD.
template <int ST, bool B>
__global__ void __launch_bounds__(256) k_update4_xcd(LatticeArgs a, FieldPtrs f, int z_base, int gx, int gy, int n) {
    const int per = (n + 7) / 8;
    const int t = (static_cast<int>(blockIdx.x) % 8) * per + static_cast<int>(blockIdx.x) / 8;
    if (t >= n) return;
    const int bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int x0 = (bx * 64 + static_cast<int>(threadIdx.x) % 64) * 4;
    const int y = by * 4 + static_cast<int>(threadIdx.x) / 64, z = z_base + bz;
    if (x0 >= a.nx || y >= a.ny) return;
    if (B) b_quad<ST>(a, f, x0, y, z);
    else e_quad<ST>(a, f, x0, y, z);
}

template <int ST>
hipError_t launch_st_xcd(const LatticeArgs &a, const FieldPtrs &f, int z0, int nzs, hipStream_t s) {
    const int gx = (a.nx / 4 + 63) / 64, gy = (a.ny + 3) / 4;
    const int n = gx * gy * nzs;
    const int blocks = (n + 7) / 8 * 8;
    k_update4_xcd<ST, true><<<blocks, 256, 0, s>>>(a, f, z0, gx, gy, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_update4_xcd<ST, false><<<blocks, 256, 0, s>>>(a, f, z0, gx, gy, n);
    return hipGetLastError();
}

template <int ST>
hipError_t launch_st(const LatticeArgs &a, const FieldPtrs &f, int z0, int nzs, hipStream_t s) {
    dim3 block(64, 4, 1), grid((a.nx / 4 + 63) / 64, (a.ny + 3) / 4, nzs);
    const size_t lds = static_cast<size_t>(4 * coef_entries(a, 1) * kCoefStride) * sizeof(float);
    k_update_b4<ST><<<grid, block, lds, s>>>(a, f, z0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.pml_materials)
        k_update_e4<ST, true><<<grid, block, lds, s>>>(a, f, z0);
    else
        k_update_e4<ST, false><<<grid, block, lds, s>>>(a, f, z0);
    return hipGetLastError();
}

hipError_t launch_range(const LatticeArgs &a, const FieldPtrs &f, int z0, int nzs, int streaming, hipStream_t s) {
    if (streaming == 6) return launch_st_xcd<2>(a, f, z0, nzs, s);
    if (streaming == 4) return launch_st_xcd<0>(a, f, z0, nzs, s);
    if (streaming == 1) return launch_st<1>(a, f, z0, nzs, s);
    if (streaming == 2) return launch_st<2>(a, f, z0, nzs, s);
    if (streaming == 3) return launch_st<3>(a, f, z0, nzs, s);
    return launch_st<0>(a, f, z0, nzs, s);
}

}  // namespace

int yee_two_pass4_supported(const LatticeArgs &a) { return a.nx % 4 == 0 ? 1 : 0; }

hipError_t launch_yee_two_pass4(const LatticeArgs &a, const FieldPtrs &f, int streaming, hipStream_t s) {
    if (!yee_two_pass4_supported(a)) return hipErrorInvalidValue;
    return launch_range(a, f, 0, a.nz, streaming, s);
}

// Slab-interleaved order: B then E on z-slab 0, then on slab 1, ... A valid leapfrog schedule:
// E on slab k needs B^{n+1/2} on slab k and on the plane below it (slab k-1, already done), and
// B on slab k+1 needs E^n on slab k+1 and the plane above it, which no earlier E pass has
// touched. The E pass of a slab re-reads what the B pass just read and wrote, so with a slab of
// a few tens of MB those reads are served by the memory-side cache instead of HBM.
hipError_t launch_yee_two_pass4_slabs(const LatticeArgs &a, const FieldPtrs &f, int slab, hipStream_t s) {
    if (!yee_two_pass4_supported(a)) return hipErrorInvalidValue;
    if (slab < 1) slab = 1;
    for (int z0 = 0; z0 < a.nz; z0 += slab) {
        const int nzs = a.nz - z0 < slab ? a.nz - z0 : slab;
        const hipError_t e = launch_range(a, f, z0, nzs, 0, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace emfd
