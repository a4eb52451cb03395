// Deterministic single-precision exponential used by every absorber / material / footprint
// coefficient of the solver, on the host and on the device.
//
// Method: Cody-Waite argument reduction x = n ln2 + r with |r| <= ln2/2 (ln2 split into a
// 16-bit head, so n * head is exact, and a tail), exp(r) by its Taylor series to degree 7 in
// Horner form (truncation error < 6e-9 relative on the reduced range), then an exact scaling
// by 2^n. Every operation is a plain IEEE float operation (the library is built without
// contraction), so the host and the GPU produce the same bits. Accuracy is within 2 ulp; the
// reference's GLSL exp carries no tighter guarantee.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace emfd {

__host__ __device__ inline float det_expf(float x) {
    if (!(x == x)) return x;               // NaN
    if (x > 88.72283935546875f) return HUGE_VALF;
    if (x < -103.972084045410156f) return 0.0f;
    const float n = rintf(x * 1.44269502162933349609f);
    const float r = (x - n * 0.693145751953125f) - n * 1.42860682030941723212e-6f;
    float p = 1.0f / 5040.0f;
    p = 1.0f / 720.0f + r * p;
    p = 1.0f / 120.0f + r * p;
    p = 1.0f / 24.0f + r * p;
    p = 1.0f / 6.0f + r * p;
    p = 0.5f + r * p;
    p = 1.0f + r * p;
    p = 1.0f + r * p;
    return ldexpf(p, static_cast<int>(n));
}

}  // namespace emfd
