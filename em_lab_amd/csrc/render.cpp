// Volume view entry points: emfd_grid_field_view hands out the grid's device planes as they are,
// emfd_render_volume runs the raymarch kernel (render.hip) over them with the viewer's set-up
// (src/VolumeRenderer.cpp:70-121: world box, cutaway clamps, step clamp).
#include "context.h"

#include <algorithm>
#include <cmath>

using namespace emfd;

extern "C" {

int emfd_grid_field_view(emfd_ctx *ctx, emfd_field_view *out) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "emfd_grid_field_view: null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "emfd_grid_field_view: grid not configured");
    const FieldPtrs f = ctx->fields(ctx->cur);
    emfd_field_view v{};
    for (int a = 0; a < 3; ++a) {
        v.size[a] = ctx->cfg.size[a];
        v.e[a] = f.e[a];
        v.b[a] = f.b[a];
    }
    v.storage = ctx->storage;
    for (int c = 0; c < 6; ++c) v.scale[c] = f.scale[c];
    v.sigma = ctx->sigma_cell.as<float>();
    v.eps_r = ctx->eps_cell.as<float>();
    v.device = ctx->device;
    v.step_count = ctx->step_count;
    *out = v;
    return EMFD_OK;
}

void emfd_render_default_settings(emfd_render_settings *out) {
    if (!out) return;
    emfd_render_settings s{};
    s.mode = EMFD_RENDER_FIELDS;
    s.signed_axis = 2;
    s.gain_e = 1.0f;
    s.gain_b = 1.0f;
    s.density = 1.0f;
    s.emission = 1.0f;
    s.step_voxels = 0.75f;
    s.material_opacity = 0.45f;
    s.exposure = 1.25f;
    for (int a = 0; a < 3; ++a) {
        s.clip_min[a] = 0.0f;
        s.clip_max[a] = 1.0f;
    }
    *out = s;
}

int emfd_render_volume(emfd_ctx *ctx, const emfd_render_settings *settings, const float *inv_view_proj,
                       const float *cam_pos, int width, int height, int format, void *image) {
    if (!ctx || !settings || !inv_view_proj || !cam_pos || !image)
        return fail(EMFD_ERR_INVALID, "emfd_render_volume: null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "emfd_render_volume: grid not configured");
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384)
        return fail(EMFD_ERR_INVALID, "emfd_render_volume: image size outside 1 .. 16384");
    const int kind = format & 0xff;
    const bool on_device = (format & EMFD_IMAGE_DEVICE) != 0;
    if ((kind != EMFD_IMAGE_RGBA8 && kind != EMFD_IMAGE_RGBA32F) || (format & ~(0xff | EMFD_IMAGE_DEVICE)))
        return fail(EMFD_ERR_INVALID, "emfd_render_volume: unknown image format");
    if (settings->mode < EMFD_RENDER_FIELDS || settings->mode > EMFD_RENDER_SIGNED_E)
        return fail(EMFD_ERR_INVALID, "emfd_render_volume: unknown mode");
    bool finite = std::isfinite(settings->gain_e) && std::isfinite(settings->gain_b) && std::isfinite(settings->density) &&
                  std::isfinite(settings->emission) && std::isfinite(settings->step_voxels) &&
                  std::isfinite(settings->material_opacity) && std::isfinite(settings->exposure);
    for (int i = 0; i < 16; ++i) finite = finite && std::isfinite(inv_view_proj[i]);
    for (int a = 0; a < 3; ++a)
        finite = finite && std::isfinite(cam_pos[a]) && std::isfinite(settings->clip_min[a]) &&
                 std::isfinite(settings->clip_max[a]);
    if (!finite) return fail(EMFD_ERR_INVALID, "emfd_render_volume: non-finite setting, matrix or eye");
    RenderArgs r{};
    for (int i = 0; i < 16; ++i) r.inv_vp[i] = inv_view_proj[i];
    // src/VolumeRenderer.cpp:44-54 (world box: longest axis 2 units, centred), :89-112 (clamps)
    const int *g = ctx->cfg.size;
    const float ws = 2.0f / static_cast<float>(std::max({g[0], g[1], g[2], 1}));
    for (int a = 0; a < 3; ++a) {
        r.cam[a] = cam_pos[a];
        r.n[a] = g[a];
        r.vol_min[a] = -0.5f * static_cast<float>(g[a]) * ws;
        r.vol_size[a] = static_cast<float>(g[a]) * ws;
        const float lo = std::clamp(std::min(settings->clip_min[a], settings->clip_max[a]), 0.0f, 1.0f);
        const float hi = std::clamp(std::max(settings->clip_min[a], settings->clip_max[a]), 0.0f, 1.0f);
        r.clip_min[a] = r.vol_min[a] + lo * r.vol_size[a];
        r.clip_max[a] = r.vol_min[a] + hi * r.vol_size[a];
    }
    r.step_voxels = std::clamp(settings->step_voxels, 0.1f, 4.0f);
    r.gain_e = settings->gain_e;
    r.gain_b = settings->gain_b;
    r.density = settings->density;
    r.emission = settings->emission;
    r.material_opacity = settings->material_opacity;
    r.exposure = settings->exposure;
    r.mode = settings->mode;
    r.signed_axis = std::clamp(settings->signed_axis, 0, 2);
    r.width = width;
    r.height = height;
    EMFD_HIP(hipSetDevice(ctx->device));
    const size_t bytes = static_cast<size_t>(width) * height * (kind == EMFD_IMAGE_RGBA8 ? 4 : 16);
    void *dst = image;
    if (!on_device) {
        // the staging image stays allocated for the next frame (a viewer renders every frame);
        // earlier work on the stream may still read a smaller one, so wait before it grows
        if (ctx->d_render.bytes < bytes) EMFD_HIP(hipStreamSynchronize(ctx->stream));
        if (int rc = ctx->d_render.ensure(bytes)) return rc;
        dst = ctx->d_render.ptr;
    }
    hipError_t e = launch_raymarch(r, ctx->fields(ctx->cur), ctx->sigma_cell.as<float>(), ctx->eps_cell.as<float>(), dst,
                                   kind == EMFD_IMAGE_RGBA8 ? 1 : 0, ctx->stream);
    if (e == hipSuccess && !on_device) e = hipMemcpyAsync(image, dst, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && !on_device) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "emfd_render_volume");
    return EMFD_OK;
}

}  // extern "C"
