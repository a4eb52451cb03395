// Radar back-projection image (RadarMonitor::computeImage, src/Radar.cpp:622-658): for every image
// point (range, azimuth, elevation) around the transmit centre, the coherent sum over the receive
// elements of their compressed, decimated echoes at the point's round-trip delay:
//   P = tx + r dir(az, el); t_i = txDelay(P) + |P - element_i|;
//   sum = sum_i w_i z_i(t_i) e^{+j w t_i} (linear interpolation between decimated samples);
//   power = |sum|^2, or -1 where P lies outside the simulated volume.
// txDelay (src/Radar.cpp:86-104) is the group delay at the carrier of the transmitting elements'
// near-field array factor, sum_j (a_j g_j / d_j) e^{-j w d_j} with g_j = max(cos(off boresight), 0).
//
// One thread per image point; element and transmitter tables in LDS. Positions and the direction
// are float and the delays and sums double, as the reference evaluates them; the compressed
// samples z are complex float. FP64 arithmetic dominates (two sincos per element and
// transmitter): the kernel is bound by the FP64 pipe, not by memory (z and the tables stay in
// L2 / LDS).
#include "internal.h"

#include <hip/hip_runtime.h>

namespace emfd {
namespace {

constexpr int kRadarBlock = 256;

struct RadarArgs {
    float tx[3];
    float axes[9];
    float lo[3], hi[3];
    double omega;       // 2 pi / wavelength
    double dt, sample_dt;
    float r0, dr, az0, daz, el0, del;
    int nr, naz, nel;
    int n_el, n_tx;
    long m;             // decimated samples per element
};

__global__ void __launch_bounds__(kRadarBlock) k_radar_backproject(RadarArgs a, const float2 *__restrict__ z,
                                                                    const float *__restrict__ el_pos,
                                                                    const double *__restrict__ el_w,
                                                                    const float *__restrict__ tx_pos,
                                                                    const double *__restrict__ tx_amp,
                                                                    double *__restrict__ power) {
    extern __shared__ double sh[];
    double *w_s = sh;                         // [n_el]
    double *amp_s = w_s + a.n_el;             // [n_tx]
    float *pos_s = reinterpret_cast<float *>(amp_s + a.n_tx);  // [n_el][3]
    float *txp_s = pos_s + 3 * a.n_el;                          // [n_tx][3]
    for (int i = threadIdx.x; i < a.n_el; i += blockDim.x) {
        w_s[i] = el_w[i];
        for (int k = 0; k < 3; ++k) pos_s[3 * i + k] = el_pos[3 * i + k];
    }
    for (int j = threadIdx.x; j < a.n_tx; j += blockDim.x) {
        amp_s[j] = tx_amp[j];
        for (int k = 0; k < 3; ++k) txp_s[3 * j + k] = tx_pos[3 * j + k];
    }
    __syncthreads();
    const long total = static_cast<long>(a.nr) * a.nel * a.naz;
    const long k = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= total) return;
    // power[(ir * nel + ie) * naz + ia]
    const int ia = static_cast<int>(k % a.naz);
    const int ie = static_cast<int>((k / a.naz) % a.nel);
    const int ir = static_cast<int>(k / (static_cast<long>(a.naz) * a.nel));

    // direction(az, el) (src/Radar.cpp:67-70), float as glm evaluates it
    const float az = a.az0 + static_cast<float>(ia) * a.daz, el = a.el0 + static_cast<float>(ie) * a.del;
    const float ar = az * 0.01745329251994329577f, er = el * 0.01745329251994329577f;
    const float sa = sinf(ar), ca = cosf(ar), se = sinf(er), ce = cosf(er);
    float dir[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) dir[c] = ((sa * ce) * a.axes[c] + se * a.axes[3 + c]) + (ca * ce) * a.axes[6 + c];
    const float r = a.r0 + static_cast<float>(ir) * a.dr;
    float P[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) P[c] = a.tx[c] + r * dir[c];
    if (P[0] < a.lo[0] || P[1] < a.lo[1] || P[2] < a.lo[2] || P[0] > a.hi[0] || P[1] > a.hi[1] || P[2] > a.hi[2]) {
        power[k] = -1.0;
        return;
    }
    // txDelay(P) (src/Radar.cpp:86-104)
    double are = 0.0, aim = 0.0, adre = 0.0, adim = 0.0, mag = 0.0;
    for (int j = 0; j < a.n_tx; ++j) {
        const float v0 = P[0] - txp_s[3 * j], v1 = P[1] - txp_s[3 * j + 1], v2 = P[2] - txp_s[3 * j + 2];
        const float len = sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
        const double d = fmax(static_cast<double>(len), 0.5);
        const float dotn = (v0 * a.axes[6] + v1 * a.axes[7]) + v2 * a.axes[8];
        const double g = fmax(static_cast<double>(dotn) / d, 0.0);
        const double s = amp_s[j] * g / d;
        double sn, cs;
        sincos(-a.omega * d, &sn, &cs);
        const double tre = s * cs, tim = s * sn;
        are += tre;
        aim += tim;
        adre += d * tre;
        adim += d * tim;
        mag += s;
    }
    double outbound;
    if (sqrt(are * are + aim * aim) < 1e-3 * mag) {
        const float u0 = P[0] - a.tx[0], u1 = P[1] - a.tx[1], u2 = P[2] - a.tx[2];
        outbound = static_cast<double>(sqrtf((u0 * u0 + u1 * u1) + u2 * u2));
    } else {
        // real(ad / a)
        outbound = (adre * are + adim * aim) / (are * are + aim * aim);
    }
    double sre = 0.0, sim = 0.0;
    for (int i = 0; i < a.n_el; ++i) {
        const float v0 = P[0] - pos_s[3 * i], v1 = P[1] - pos_s[3 * i + 1], v2 = P[2] - pos_s[3 * i + 2];
        const double t = outbound + static_cast<double>(sqrtf((v0 * v0 + v1 * v1) + v2 * v2));
        const double s = (t + 0.5 * a.dt) / a.sample_dt;
        if (s < 0.0 || s >= static_cast<double>(a.m - 1)) continue;
        const long m0 = static_cast<long>(s);
        const float f = static_cast<float>(s - static_cast<double>(m0));
        const float2 z0 = z[m0 * a.n_el + i], z1 = z[(m0 + 1) * a.n_el + i];
        const float vre = z0.x * (1.0f - f) + z1.x * f, vim = z0.y * (1.0f - f) + z1.y * f;
        double sn, cs;
        sincos(a.omega * t, &sn, &cs);
        const double wre = w_s[i] * static_cast<double>(vre), wim = w_s[i] * static_cast<double>(vim);
        sre += wre * cs - wim * sn;
        sim += wre * sn + wim * cs;
    }
    power[k] = sre * sre + sim * sim;
}

}  // namespace

hipError_t launch_radar_backproject(const emfd_radar_image_setup &s, const float *z, long m, int n_el,
                                    const float *el_pos, const double *el_w, const float *tx_pos,
                                    const double *tx_amp, int n_tx, double *power, hipStream_t st) {
    RadarArgs a{};
    for (int c = 0; c < 3; ++c) {
        a.tx[c] = s.tx[c];
        a.lo[c] = s.interior_min[c];
        a.hi[c] = s.interior_max[c];
    }
    for (int c = 0; c < 9; ++c) a.axes[c] = s.axes[c];
    a.omega = 6.283185307179586476925286766559 / static_cast<double>(s.wavelength);
    a.dt = s.dt;
    a.sample_dt = s.sample_dt;
    a.r0 = s.r0;
    a.dr = s.dr;
    a.az0 = s.az0;
    a.daz = s.daz;
    a.el0 = s.el0;
    a.del = s.del;
    a.nr = s.nr;
    a.naz = s.naz;
    a.nel = s.nel;
    a.n_el = n_el;
    a.n_tx = n_tx;
    a.m = m;
    const long total = static_cast<long>(s.nr) * s.nel * s.naz;
    const size_t lds = static_cast<size_t>(n_el + n_tx) * sizeof(double) + static_cast<size_t>(3 * (n_el + n_tx)) * sizeof(float);
    const long blocks = (total + kRadarBlock - 1) / kRadarBlock;
    k_radar_backproject<<<static_cast<unsigned>(blocks), kRadarBlock, lds, st>>>(a, reinterpret_cast<const float2 *>(z), el_pos,
                                                                             el_w, tx_pos, tx_amp, power);
    return hipGetLastError();
}

}  // namespace emfd
