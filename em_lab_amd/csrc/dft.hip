// Running discrete Fourier transforms in FP64, next to the reference's FP32 exponential means
// (no counterpart in the reference, which keeps only the two-stage EMA of port_accumulate.comp and
// ntff_accumulate.comp and transforms the recorded FP32 port series on the host).
//
//   X(w) = sum_n x_n e^{-i w t_n},  t_n = n dt for E samples, (n - 1/2) dt for B / H samples,
//
// with x_n the FP32 sample of the step (the port gap V and I, the Huygens-node E and H exactly as
// the EMA path forms them) and the rotations e^{-i w t_n} computed in FP64 on the host per step.
// A pulse run long enough to ring down then gives the spectrum at each wavelength, where the EMA
// gives a time-windowed value.
#include "port_device.h"

namespace emfd {
namespace {

// One thread per (port, frequency): the port's gap sample (as port_sample forms it) into its FP64
// accumulators [V re, V im, I re, I im]. rot: this step's [frequency][cos wtE, sin wtE, cos wtH,
// sin wtH].
__global__ void __launch_bounds__(64) k_port_dft(LatticeArgs a, FieldPtrs f, const PortDev *ports, int n_ports,
                                                 const double *rot, int n_freq, double *acc) {
    const int t = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x);
    if (t >= n_ports * n_freq) return;
    const int p = t / n_freq, q = t % n_freq;
    float V, I;
    port_gap(a, f, ports[p], &V, &I);
    const double *r = rot + 4 * q;
    double *x = acc + 4 * static_cast<size_t>(t);
    x[0] += static_cast<double>(V) * r[0];
    x[1] -= static_cast<double>(V) * r[1];
    x[2] += static_cast<double>(I) * r[2];
    x[3] -= static_cast<double>(I) * r[3];
}

// The Huygens points' FP64 accumulators over a run of K sampled steps (the deferred far-field
// samples: slot k holds E x y z then H x y z at every point). Layout: per frequency 12 planes of
// n_points doubles (E re x y z, E im, H re, H im). rot: step k, frequency q at
// rot[k * stride + 4 q].
template <int K>
__global__ void __launch_bounds__(256) k_ntff_dft(NtffBox box, const double *rot, int stride, int n_freq,
                                                  const float *samples, double *acc) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= box.n_points) return;
    const size_t np = static_cast<size_t>(box.n_points);
    float smp[K][6];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) smp[k][c] = samples[(6 * k + c) * np + i];
    for (int q = 0; q < n_freq; ++q) {
        double *x = acc + static_cast<size_t>(12 * q) * np + i;
        double v[12];
#pragma unroll
        for (int j = 0; j < 12; ++j) v[j] = x[j * np];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double *r = rot + static_cast<size_t>(k) * stride + 4 * q;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double e = smp[k][c], h = smp[k][3 + c];
                v[c] += e * r[0];
                v[3 + c] -= e * r[1];
                v[6 + c] += h * r[2];
                v[9 + c] -= h * r[3];
            }
        }
#pragma unroll
        for (int j = 0; j < 12; ++j) x[j * np] = v[j];
    }
}

// FP64 accumulators of one frequency -> the FP32 stage-2 phasor planes (12 + 3 q + c: E re, E im,
// H re, H im) that the radiation integrals read; stage-1 planes are left as they are.
__global__ void __launch_bounds__(256) k_ntff_dft_planes(const double *acc, int n_points, float *ph) {
    const int i = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (i >= n_points) return;
    const size_t np = static_cast<size_t>(n_points);
#pragma unroll
    for (int j = 0; j < 12; ++j) ph[(12 + j) * np + i] = static_cast<float>(acc[j * np + i]);
}

template <int K>
hipError_t launch_dft_k(const NtffBox &box, const double *rot, int stride, int n_freq, const float *samples,
                        double *acc, hipStream_t s) {
    k_ntff_dft<K><<<(box.n_points + 255) / 256, 256, 0, s>>>(box, rot, stride, n_freq, samples, acc);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_port_dft(const LatticeArgs &a, const FieldPtrs &f, const PortDev *ports, int n_ports,
                           const double *rot, int n_freq, double *acc, hipStream_t s) {
    const int n = n_ports * n_freq;
    if (n <= 0) return hipSuccess;
    k_port_dft<<<(n + 63) / 64, 64, 0, s>>>(a, f, ports, n_ports, rot, n_freq, acc);
    return hipGetLastError();
}

hipError_t launch_ntff_dft(const NtffBox &box, const double *rot, int stride, int steps, int n_freq,
                           const float *samples, double *acc, hipStream_t s) {
    if (box.n_points <= 0 || steps <= 0 || n_freq <= 0) return hipSuccess;
    switch (steps) {
    case 1: return launch_dft_k<1>(box, rot, stride, n_freq, samples, acc, s);
    case 2: return launch_dft_k<2>(box, rot, stride, n_freq, samples, acc, s);
    case 3: return launch_dft_k<3>(box, rot, stride, n_freq, samples, acc, s);
    case 4: return launch_dft_k<4>(box, rot, stride, n_freq, samples, acc, s);
    case 5: return launch_dft_k<5>(box, rot, stride, n_freq, samples, acc, s);
    case 6: return launch_dft_k<6>(box, rot, stride, n_freq, samples, acc, s);
    case 7: return launch_dft_k<7>(box, rot, stride, n_freq, samples, acc, s);
    case 8: return launch_dft_k<8>(box, rot, stride, n_freq, samples, acc, s);
    default:
        // runs longer than 8 steps: two passes (the sums are the same up to FP64 addition order)
        if (hipError_t e = launch_dft_k<8>(box, rot, stride, n_freq, samples, acc, s); e != hipSuccess) return e;
        return launch_ntff_dft(box, rot + static_cast<size_t>(8) * stride, stride, steps - 8, n_freq,
                               samples + static_cast<size_t>(8) * 6 * box.n_points, acc, s);
    }
}

hipError_t launch_ntff_dft_planes(const double *acc, int n_points, float *ph, hipStream_t s) {
    if (n_points <= 0) return hipSuccess;
    k_ntff_dft_planes<<<(n_points + 255) / 256, 256, 0, s>>>(acc, n_points, ph);
    return hipGetLastError();
}

}  // namespace emfd
