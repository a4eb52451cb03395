// FP64 stepper: the same leapfrog, absorber and material model as the FP32 kernels, with every
// field value and every coefficient in double precision (profiles, binary16 edge materials and
// wire factors are exact in double). It is the accuracy reference of the reduced-precision
// storage: exp and phi(x) = -expm1(-x) / x come from the device's double-precision library,
// the source signals are evaluated in double on the device. One thread per cell.
#include "port_device.h"
#include "signal.h"
#include "yee_device.h"

namespace emfd {
namespace {

__device__ __forceinline__ double phi_d(double x) { return x > 0.0 ? -expm1(-x) / x : 1.0; }

__device__ __forceinline__ double prof_d(const LatticeArgs &a, int x, int hx, int y, int hy, int z, int hz) {
    const int s = a.prof_stride;
    return static_cast<double>(a.prof[(2 * 0 + hx) * s + x]) + static_cast<double>(a.prof[(2 * 1 + hy) * s + y]) +
           static_cast<double>(a.prof[(2 * 2 + hz) * s + z]);
}

__device__ __forceinline__ uint32_t id_at(const LatticeArgs &a, size_t p) { return a.mat_id ? load_id(a, p) : 0u; }

// Ca, Cb of the three E edges of texel (x, y, z) in double
__device__ __forceinline__ void edge_coefficients_d(const LatticeArgs &a, int x, int y, int z, uint32_t id,
                                                    double ca[3], double cb[3]) {
    double sig[3] = {0.0, 0.0, 0.0}, eps[3] = {1.0, 1.0, 1.0};
    if (id != 0u) {
        const float *t = a.table + static_cast<size_t>(id) * kTableStride;
        for (int c = 0; c < 3; ++c) {
            sig[c] = t[kSig + c];
            eps[c] = t[kEps + c];
        }
    }
    sig[0] += prof_d(a, x, 1, y, 0, z, 0);
    sig[1] += prof_d(a, x, 0, y, 1, z, 0);
    sig[2] += prof_d(a, x, 0, y, 0, z, 1);
    const double dt = a.dt;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double dte = dt / eps[c];
        const double xx = sig[c] * dte;
        ca[c] = exp(-xx);
        cb[c] = dte * phi_d(xx);
    }
}

__global__ void __launch_bounds__(256) k_update_b_f64(LatticeArgs a, FieldPtrs f) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= a.nx || y >= a.ny) return;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x;
    const double *E[3] = {reinterpret_cast<const double *>(f.e[0]), reinterpret_cast<const double *>(f.e[1]),
                          reinterpret_cast<const double *>(f.e[2])};
    double *B[3] = {reinterpret_cast<double *>(f.b[0]), reinterpret_cast<double *>(f.b[1]),
                    reinterpret_cast<double *>(f.b[2])};
    double e0[3], ex[3], ey[3], ez[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e0[c] = E[c][p];
        ex[c] = x + 1 < a.nx ? E[c][p + 1] : 0.0;
        ey[c] = y + 1 < a.ny ? E[c][p + sy] : 0.0;
        ez[c] = z + 1 < a.nz ? E[c][p + sz] : 0.0;
    }
    double t0[3] = {ey[2] - e0[2], ez[0] - e0[0], ex[1] - e0[1]};
    double t1[3] = {ez[1] - e0[1], ex[2] - e0[2], ey[0] - e0[0]};
    const uint32_t id = id_at(a, p);
    if (a.has_wire && id != 0u) {
        const float *w = a.table + static_cast<size_t>(id) * kTableStride + kWire;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            t0[c] *= static_cast<double>(w[2 * c]);
            t1[c] *= static_cast<double>(w[2 * c + 1]);
        }
    }
    const double dt = a.dt;
    const double s[3] = {prof_d(a, x, 0, y, 1, z, 1) * dt, prof_d(a, x, 1, y, 0, z, 1) * dt,
                         prof_d(a, x, 1, y, 1, z, 0) * dt};
#pragma unroll
    for (int c = 0; c < 3; ++c) B[c][p] = exp(-s[c]) * B[c][p] - dt * phi_d(s[c]) * (t0[c] - t1[c]);
}

__global__ void __launch_bounds__(256) k_update_e_f64(LatticeArgs a, FieldPtrs f) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= a.nx || y >= a.ny) return;
    const size_t sy = a.nx, sz = static_cast<size_t>(a.nx) * a.ny;
    const size_t p = z * sz + y * sy + x;
    const double *B[3] = {reinterpret_cast<const double *>(f.b[0]), reinterpret_cast<const double *>(f.b[1]),
                          reinterpret_cast<const double *>(f.b[2])};
    double *E[3] = {reinterpret_cast<double *>(f.e[0]), reinterpret_cast<double *>(f.e[1]),
                    reinterpret_cast<double *>(f.e[2])};
    double b0[3], bx[3], by[3], bz[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b0[c] = B[c][p];
        bx[c] = x > 0 ? B[c][p - 1] : 0.0;
        by[c] = y > 0 ? B[c][p - sy] : 0.0;
        bz[c] = z > 0 ? B[c][p - sz] : 0.0;
    }
    const double curl[3] = {(b0[2] - by[2]) - (b0[1] - bz[1]), (b0[0] - bz[0]) - (b0[2] - bx[2]),
                            (b0[1] - bx[1]) - (b0[0] - by[0])};
    double ca[3], cb[3];
    edge_coefficients_d(a, x, y, z, id_at(a, p), ca, cb);
#pragma unroll
    for (int c = 0; c < 3; ++c) E[c][p] = ca[c] * E[c][p] + cb[c] * curl[c];
}

__device__ __forceinline__ double footprint_d(const float *s, double x, double y, double z) {
    const double dx = fmax(fabs(x - s[0]) - s[4], 0.0) / s[12];
    const double dy = fmax(fabs(y - s[1]) - s[5], 0.0) / s[13];
    const double dz = fmax(fabs(z - s[2]) - s[6], 0.0) / s[14];
    return exp(-(dx * dx + dy * dy + dz * dz));
}

// apply_sources (shaders/apply_sources.comp:37-79) in double: signals at t from the source
// descriptions, footprints and Cb in double
__global__ void __launch_bounds__(64) k_sources_f64(LatticeArgs a, FieldPtrs f, const int *cells, int n_cells,
                                                    const int *lists, const float *packed,
                                                    const emfd_source_desc *descs, double t) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_cells) return;
    const int x = cells[4 * id + 0], y = cells[4 * id + 1], z = cells[4 * id + 2];
    const int first = cells[4 * id + 3], end = cells[4 * (id + 1) + 3];
    double J[3] = {0.0, 0.0, 0.0};
    for (int k = first; k < end; ++k) {
        const int i = lists[k];
        const double signal = source_signal_d(descs[i], t);
        if (signal == 0.0) continue;
        const float *s = packed + 16 * i;
        const double w[3] = {footprint_d(s, x + 0.5, y, z), footprint_d(s, x, y + 0.5, z), footprint_d(s, x, y, z + 0.5)};
        const double as = static_cast<double>(s[11]) * signal;
#pragma unroll
        for (int c = 0; c < 3; ++c) J[c] += static_cast<double>(s[8 + c]) * as * w[c];
    }
    if (J[0] == 0.0 && J[1] == 0.0 && J[2] == 0.0) return;
    const size_t p = (static_cast<size_t>(z) * a.ny + y) * a.nx + x;
    double ca[3], cb[3];
    edge_coefficients_d(a, x, y, z, id_at(a, p), ca, cb);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double *Ec = reinterpret_cast<double *>(f.e[c]);
        Ec[p] = Ec[p] - cb[c] * J[c];
    }
}

}  // namespace

hipError_t launch_yee_f64(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s) {
    dim3 block(64, 4, 1), grid((a.nx + 63) / 64, (a.ny + 3) / 4, a.nz);
    k_update_b_f64<<<grid, block, 0, s>>>(a, f);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_update_e_f64<<<grid, block, 0, s>>>(a, f);
    return hipGetLastError();
}

hipError_t launch_sources_f64(const LatticeArgs &a, const FieldPtrs &f, const int *cells, int n_cells,
                              const int *lists, const float *packed, const emfd_source_desc *descs, double t,
                              hipStream_t s) {
    if (n_cells <= 0) return hipSuccess;
    k_sources_f64<<<(n_cells + 63) / 64, 64, 0, s>>>(a, f, cells, n_cells, lists, packed, descs, t);
    return hipGetLastError();
}

}  // namespace emfd
