// Radar back-projection entry point (emfd_radar_backproject): uploads the echoes and element
// tables, runs k_radar_backproject on the context's stream and reads the image back.
#include "context.h"

#include <vector>

using namespace emfd;

extern "C" {

int emfd_radar_backproject(emfd_ctx *ctx, const emfd_radar_image_setup *setup, const float *z, int64_t m, int n_el,
                           const float *element_pos, const double *weights, const float *tx_pos, const double *tx_amp,
                           int n_tx, double *power) {
    if (!ctx || !setup || !power) return fail(EMFD_ERR_INVALID, "emfd_radar_backproject: null argument");
    if (n_el <= 0 || m < 2 || !z || !element_pos || !weights)
        return fail(EMFD_ERR_INVALID, "emfd_radar_backproject: needs at least one element and two samples");
    if (n_tx < 0 || (n_tx > 0 && (!tx_pos || !tx_amp)))
        return fail(EMFD_ERR_INVALID, "emfd_radar_backproject: null transmitter table");
    if (setup->nr <= 0 || setup->naz <= 0 || setup->nel <= 0 || !(setup->sample_dt > 0.0) || !(setup->wavelength > 0.0f))
        return fail(EMFD_ERR_INVALID, "emfd_radar_backproject: empty image or non-positive sample spacing");
    if (n_el > 4096 || n_tx > 4096) return fail(EMFD_ERR_INVALID, "emfd_radar_backproject: at most 4096 elements");
    const size_t total = static_cast<size_t>(setup->nr) * setup->nel * setup->naz;
    EMFD_HIP(hipSetDevice(ctx->device));
    const size_t zb = static_cast<size_t>(m) * n_el * 2 * sizeof(float);
    const size_t tb = static_cast<size_t>(n_el) * (3 * sizeof(float) + sizeof(double)) +
                      static_cast<size_t>(n_tx) * (3 * sizeof(float) + sizeof(double)) + 64;
    DevBuf dz, dt, dp;
    if (int rc = dz.ensure(zb)) return rc;
    if (int rc = dt.ensure(tb)) {
        dz.release();
        return rc;
    }
    if (int rc = dp.ensure(total * sizeof(double))) {
        dz.release();
        dt.release();
        return rc;
    }
    char *base = dt.as<char>();
    double *d_w = reinterpret_cast<double *>(base);
    double *d_amp = d_w + n_el;
    float *d_pos = reinterpret_cast<float *>(d_amp + n_tx);
    float *d_txp = d_pos + 3 * n_el;
    hipError_t e = hipMemcpyAsync(dz.ptr, z, zb, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_w, weights, n_el * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && n_tx > 0) e = hipMemcpyAsync(d_amp, tx_amp, n_tx * sizeof(double), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pos, element_pos, 3 * n_el * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && n_tx > 0) e = hipMemcpyAsync(d_txp, tx_pos, 3 * n_tx * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = launch_radar_backproject(*setup, dz.as<float>(), static_cast<long>(m), n_el, d_pos, d_w, d_txp, d_amp, n_tx,
                                     dp.as<double>(), ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(power, dp.ptr, total * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dz.release();
    dt.release();
    dp.release();
    if (e != hipSuccess) return hip_fail(e, "emfd_radar_backproject");
    return EMFD_OK;
}

}  // extern "C"
