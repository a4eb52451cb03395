/*
 * CPU oracle of EM_Lab's volume raymarch (shaders/raymarch.frag:38-198, set up by
 * src/VolumeRenderer.cpp:44-54,70-121).
 *
 * TEST INFRASTRUCTURE ONLY: a scalar C restatement that checks the HIP raymarch; the product
 * library never calls it. Textures are the reference's RGBA32F texels (E with sigma in alpha,
 * B with eps_r in alpha), sampled as GL_LINEAR with GL_CLAMP_TO_EDGE over texel centres
 * (src/SimulationGrid.cpp:167-173). Floating point follows the shader in float, one fixed
 * order per expression (built with -ffp-contract=off).
 */
#include "oracle.h"

#include <math.h>

typedef struct {
    float x, y, z;
} rv3;

static rv3 v3(float x, float y, float z) {
    rv3 r = {x, y, z};
    return r;
}
static rv3 add(rv3 a, rv3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static rv3 sub(rv3 a, rv3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static rv3 scl(rv3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
static float dot(rv3 a, rv3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static float len(rv3 a) { return sqrtf(dot(a, a)); }
static rv3 norm(rv3 a) {
    float l = len(a);
    return v3(a.x / l, a.y / l, a.z / l);
}
static float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static rv3 mixv(rv3 x, rv3 y, float a) { return v3(mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)); }
static float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static float fractf(float x) { return x - floorf(x); }

/* raymarch.frag:49-51 */
float or_gradient_noise(float px, float py) {
    return fractf(52.9829189f * fractf(px * 0.06711056f + py * 0.00583715f));
}
/* raymarch.frag:60 */
static float response(float x) {
    float x2 = x * x;
    return x2 / (1.0f + x2);
}

/* texture(): channel ch of the RGBA texels at normalised coordinates (u, v, w) */
float or_texture(const float *tex, int nx, int ny, int nz, int ch, float u, float v, float w) {
    const int n[3] = {nx, ny, nz};
    const float c[3] = {u, v, w};
    int lo[3], hi[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
        float s = c[a] * (float)n[a] - 0.5f;
        float fl = floorf(s);
        int i = (int)fminf(fmaxf(fl, -1.0f), (float)n[a]);
        f[a] = s - fl;
        lo[a] = i < 0 ? 0 : (i > n[a] - 1 ? n[a] - 1 : i);
        hi[a] = i + 1 < 0 ? 0 : (i + 1 > n[a] - 1 ? n[a] - 1 : i + 1);
    }
    float t[8];
    for (int k = 0; k < 8; ++k) {
        int x = (k & 1) ? hi[0] : lo[0], y = (k & 2) ? hi[1] : lo[1], z = (k & 4) ? hi[2] : lo[2];
        t[k] = tex[(((size_t)z * (size_t)ny + (size_t)y) * (size_t)nx + (size_t)x) * 4 + (size_t)ch];
    }
    float a0 = mixf(t[0], t[1], f[0]), a1 = mixf(t[2], t[3], f[0]);
    float b0 = mixf(t[4], t[5], f[0]), b1 = mixf(t[6], t[7], f[0]);
    return mixf(mixf(a0, a1, f[1]), mixf(b0, b1, f[1]), f[2]);
}

/* raymarch.frag:38-47 */
void or_box_hit(const float *ro, const float *rd, const float *bmin, const float *bmax, float *t_enter, float *t_exit) {
    float tmin[3], tmax[3];
    for (int a = 0; a < 3; ++a) {
        float d = fabsf(rd[a]) < 1e-8f ? 1e-8f : rd[a];
        float inv = 1.0f / d;
        float t0 = (bmin[a] - ro[a]) * inv, t1 = (bmax[a] - ro[a]) * inv;
        tmin[a] = fminf(t0, t1);
        tmax[a] = fmaxf(t0, t1);
    }
    *t_enter = fmaxf(fmaxf(tmin[0], tmin[1]), tmin[2]);
    *t_exit = fminf(fminf(tmax[0], tmax[1]), tmax[2]);
}

/* raymarch.frag:66-81 */
static rv3 material_normal(const or_render_args *r, const float *E, const float *B, rv3 uvw, int metal, rv3 rd) {
    const float *tex = metal ? E : B;
    const float hx = 1.0f / (float)r->n[0], hy = 1.0f / (float)r->n[1], hz = 1.0f / (float)r->n[2];
    const int nx = r->n[0], ny = r->n[1], nz = r->n[2];
    rv3 g = v3(or_texture(tex, nx, ny, nz, 3, uvw.x + hx, uvw.y, uvw.z) - or_texture(tex, nx, ny, nz, 3, uvw.x - hx, uvw.y, uvw.z),
               or_texture(tex, nx, ny, nz, 3, uvw.x, uvw.y + hy, uvw.z) - or_texture(tex, nx, ny, nz, 3, uvw.x, uvw.y - hy, uvw.z),
               or_texture(tex, nx, ny, nz, 3, uvw.x, uvw.y, uvw.z + hz) - or_texture(tex, nx, ny, nz, 3, uvw.x, uvw.y, uvw.z - hz));
    float l = len(g);
    return l > 1e-5f ? v3(-g.x / l, -g.y / l, -g.z / l) : scl(rd, -1.0f);
}

/* raymarch.frag:83-94 */
static rv3 shade_surface(rv3 n, rv3 rd, rv3 albedo, float specular) {
    rv3 key = norm(v3(0.45f, -0.35f, 0.82f));
    rv3 v = scl(rd, -1.0f);
    float ndl = fmaxf(dot(n, key), 0.0f);
    float head = fmaxf(dot(n, v), 0.0f);
    rv3 hv = norm(add(key, v));
    float spec = powf(fmaxf(dot(n, hv), 0.0f), 48.0f) * specular;
    float rim = powf(1.0f - head, 3.0f);
    float k = (0.18f + 0.55f * ndl) + 0.30f * head;
    return add(add(scl(albedo, k), v3(spec, spec, spec)), scl(v3(0.35f, 0.45f, 0.6f), rim));
}

/* one pixel (px, py), bottom row 0: raymarch.frag:103-198 */
void or_render_pixel(const or_render_args *r, const float *E, const float *B, int px, int py, float *rgba) {
    const float fx = (float)px + 0.5f, fy = (float)py + 0.5f;
    const float ndc[4] = {2.0f * fx / (float)r->width - 1.0f, 2.0f * fy / (float)r->height - 1.0f, 1.0f, 1.0f};
    float far[4];
    for (int i = 0; i < 4; ++i)
        far[i] = ((r->inv_vp[i] * ndc[0] + r->inv_vp[4 + i] * ndc[1]) + r->inv_vp[8 + i] * ndc[2]) + r->inv_vp[12 + i] * ndc[3];
    rv3 ro = v3(r->cam[0], r->cam[1], r->cam[2]);
    rv3 rd = norm(sub(v3(far[0] / far[3], far[1] / far[3], far[2] / far[3]), ro));
    float h = clampf(rd.z * 0.5f + 0.5f, 0.0f, 1.0f);
    rv3 bg = mixv(v3(0.004f, 0.006f, 0.014f), v3(0.018f, 0.028f, 0.060f), h);
    const float rof[3] = {ro.x, ro.y, ro.z}, rdf[3] = {rd.x, rd.y, rd.z};
    float t0, t1;
    or_box_hit(rof, rdf, r->clip_min, r->clip_max, &t0, &t1);
    t0 = fmaxf(t0, 0.0f);
    rv3 accum = v3(0.0f, 0.0f, 0.0f);
    float alpha = 0.0f;
    if (t0 < t1) {
        const int nx = r->n[0], ny = r->n[1], nz = r->n[2];
        float voxel = r->vol_size[0] / (float)nx;
        float ds = fmaxf(r->step_voxels, 0.05f) * voxel;
        float svox = ds / voxel;
        float t = t0 + ds * or_gradient_noise(fx, fy);
        int steps = (int)((t1 - t) / ds) + 1;
        if (steps > 4096) steps = 4096;
        for (int i = 0; i < steps; ++i, t += ds) {
            rv3 q = sub(add(ro, scl(rd, t)), v3(r->vol_min[0], r->vol_min[1], r->vol_min[2]));
            rv3 uvw = v3(q.x / r->vol_size[0], q.y / r->vol_size[1], q.z / r->vol_size[2]);
            rv3 e = v3(or_texture(E, nx, ny, nz, 0, uvw.x, uvw.y, uvw.z), or_texture(E, nx, ny, nz, 1, uvw.x, uvw.y, uvw.z),
                       or_texture(E, nx, ny, nz, 2, uvw.x, uvw.y, uvw.z));
            float sig = or_texture(E, nx, ny, nz, 3, uvw.x, uvw.y, uvw.z);
            rv3 b = v3(or_texture(B, nx, ny, nz, 0, uvw.x, uvw.y, uvw.z), or_texture(B, nx, ny, nz, 1, uvw.x, uvw.y, uvw.z),
                       or_texture(B, nx, ny, nz, 2, uvw.x, uvw.y, uvw.z));
            float eps = or_texture(B, nx, ny, nz, 3, uvw.x, uvw.y, uvw.z);

            rv3 emission = v3(0.0f, 0.0f, 0.0f);
            float extinction = 0.0f;
            if (r->mode == 0 || r->mode == 1) {
                float iE = response(len(e) * r->gain_e);
                rv3 col = mixv(v3(0.02f, 0.22f, 1.0f), v3(0.05f, 0.95f, 1.0f), iE);
                col = mixv(col, v3(0.8f, 1.0f, 1.0f), smoothstepf(0.8f, 1.0f, iE));
                emission = add(emission, scl(col, iE * iE * 2.0f));
                extinction += iE;
            }
            if (r->mode == 0 || r->mode == 2) {
                float iB = response(len(b) * r->gain_b);
                rv3 col = mixv(v3(1.0f, 0.03f, 0.10f), v3(1.0f, 0.10f, 0.62f), iB);
                col = mixv(col, v3(1.0f, 0.85f, 1.0f), smoothstepf(0.8f, 1.0f, iB));
                emission = add(emission, scl(col, iB * iB * 2.0f));
                extinction += iB;
            }
            if (r->mode == 3) {
                rv3 s = v3(e.y * b.z - e.z * b.y, e.z * b.x - e.x * b.z, e.x * b.y - e.y * b.x);
                float iS = response(len(s) * r->gain_e * r->gain_b);
                rv3 col = mixv(v3(1.0f, 0.35f, 0.02f), v3(1.0f, 0.85f, 0.25f), iS);
                col = mixv(col, v3(1.0f, 1.0f, 0.85f), smoothstepf(0.8f, 1.0f, iS));
                emission = add(emission, scl(col, iS * iS * 2.0f));
                extinction += iS;
            }
            if (r->mode == 4) {
                float comp = r->signed_axis == 0 ? e.x : (r->signed_axis == 1 ? e.y : e.z);
                float v = comp * r->gain_e;
                float iV = response(v);
                rv3 col = v >= 0.0f ? mixv(v3(0.02f, 0.3f, 1.0f), v3(0.1f, 1.0f, 1.0f), iV)
                                    : mixv(v3(0.9f, 0.15f, 0.02f), v3(1.0f, 0.65f, 0.1f), iV);
                emission = add(emission, scl(col, iV * iV * 2.0f));
                extinction += iV;
            }
            emission = scl(emission, 0.045f * r->emission);
            extinction *= 0.022f * r->density;

            float metal = smoothstepf(1.0f, 20.0f, sig);
            float glass = clampf((eps - 1.0f) * 0.5f, 0.0f, 1.0f);
            float absorber = smoothstepf(0.02f, 1.0f, sig) * (1.0f - metal);
            rv3 surface = v3(0.0f, 0.0f, 0.0f);
            float mat_ext = 0.0f;
            if ((metal + glass) + absorber > 0.01f) {
                rv3 n = material_normal(r, E, B, uvw, metal >= glass, rd);
                rv3 mc = shade_surface(n, rd, v3(0.50f, 0.53f, 0.58f), 0.9f);
                rv3 gc = shade_surface(n, rd, v3(0.30f, 0.65f, 0.70f), 0.6f);
                rv3 lc = shade_surface(n, rd, v3(0.35f, 0.22f, 0.12f), 0.1f);
                float wM = metal * r->material_opacity;
                float wG = glass * r->material_opacity * 0.02f;
                float wA = absorber * r->material_opacity * 0.4f;
                mat_ext = (wM + wG) + wA;
                rv3 sum = add(add(scl(mc, wM), scl(gc, wG)), scl(lc, wA));
                float d = fmaxf(mat_ext, 1e-6f);
                surface = v3(sum.x / d, sum.y / d, sum.z / d);
            }
            float a_field = 1.0f - expf(-extinction * svox);
            float a_mat = 1.0f - expf(-mat_ext * svox);
            float tr = 1.0f - alpha;
            accum = add(accum, scl(add(scl(scl(emission, svox), 1.0f - a_mat), scl(surface, a_mat)), tr));
            alpha += tr * (1.0f - (1.0f - a_field) * (1.0f - a_mat));
            if (alpha > 0.995f) break;
        }
    }
    rv3 c = add(accum, scl(bg, 1.0f - alpha));
    float o[3] = {c.x, c.y, c.z};
    float dither = (or_gradient_noise(fx + 17.0f, fy + 17.0f) - 0.5f) / 255.0f;
    for (int k = 0; k < 3; ++k) {
        o[k] = 1.0f - expf(-o[k] * r->exposure);
        o[k] = powf(o[k], 1.0f / 2.2f);
        o[k] += dither;
        rgba[k] = o[k];
    }
    rgba[3] = 1.0f;
}

void or_render_volume(const or_render_args *r, const float *E, const float *B, float *rgba) {
    #pragma omp parallel for schedule(dynamic, 4)
    for (int py = 0; py < r->height; ++py)
        for (int px = 0; px < r->width; ++px)
            or_render_pixel(r, E, B, px, py, rgba + ((size_t)py * (size_t)r->width + (size_t)px) * 4);
}
