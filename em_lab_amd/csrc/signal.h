// SourceDesc::signal (src/SimulationGrid.cpp:99-130) in double precision, shared by the host
// (per-step signals cast to float) and the FP64 stepper's device-side source pass.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/emfdtd.h"

namespace emfd {

__host__ __device__ inline double source_signal_d(const emfd_source_desc &s, double tAbs) {
    const double wl = fmax(static_cast<double>(s.wavelength), 1e-3);
    const double t = tAbs - fmax(static_cast<double>(s.delay_periods), 0.0) * wl;
    if (!s.enabled || t < 0.0) return 0.0;
    const double period = wl;
    const double w = 6.283185307179586476925286766559 / wl;
    if (s.waveform == 0) {
        const double rampT = fmax(static_cast<double>(s.ramp_periods), 0.0) * period;
        double ramp = 1.0;
        if (rampT > 0.0 && t < rampT) {
            const double r = t / rampT;
            ramp = r * r * r * (10.0 + r * (-15.0 + 6.0 * r));
        }
        return ramp * sin(w * t + s.phase);
    }
    const double tau = fmax(static_cast<double>(s.ramp_periods), 0.1) * period;
    const double t0 = 3.0 * tau;
    double tl = t;
    if (s.repeat_periods > 0.0f) {
        const double rep = fmax(static_cast<double>(s.repeat_periods) * period, 2.0 * t0);
        tl = fmod(t, rep);
    }
    const double u = (tl - t0) / tau;
    if (fabs(u) > 6.0) return 0.0;
    return exp(-u * u) * sin(w * (tl - t0) + s.phase);
}

}  // namespace emfd
