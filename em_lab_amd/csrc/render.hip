// Volume view of the fields: the emission / absorption raymarch EM_Lab's viewer draws over its
// field textures (shaders/raymarch.frag:107-198, set up by src/VolumeRenderer.cpp:70-121), here
// read straight from the lattice's planes (zero copy: the grid's current E and B planes and its
// per-cell sigma / eps_r, the textures' alpha channels).
//
// One thread per pixel; a wavefront holds an 8 x 8 pixel square, so its rays stay close and their
// trilinear samples share cache lines. Samples follow GL_LINEAR with GL_CLAMP_TO_EDGE on texel
// centres (src/SimulationGrid.cpp:167-173). The float operations keep one fixed order, which the
// oracle's restatement shares, so the two agree up to the last bits of exp / pow.
#include "port_device.h"

namespace emfd {
namespace {

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float len3(V3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ V3 unit(V3 a) {
    const float l = len3(a);
    return v3(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// GLSL mix (x (1 - a) + y a), clamp, smoothstep, fract
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ V3 mix3(V3 x, V3 y, float a) { return v3(mixf(x.x, y.x, a), mixf(x.y, y.y, a), mixf(x.z, y.z, a)); }
__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ float smooth(float e0, float e1, float x) {
    const float t = clamp01((x - e0) / (e1 - e0));
    return t * t * (3.0f - 2.0f * t);
}
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }
__device__ __forceinline__ float gradient_noise(float px, float py) {
    return fract(52.9829189f * fract(px * 0.06711056f + py * 0.00583715f));
}
// x^2 / (1 + x^2): quiet near zero, 1 for strong fields
__device__ __forceinline__ float response(float x) {
    const float x2 = x * x;
    return x2 / (1.0f + x2);
}

// One texture lookup of a plane set: trilinear over texel centres, indices clamped to the edge.
struct Tap {
    size_t p[8];  // corner indices, bit 0: +x, bit 1: +y, bit 2: +z
    float f[3];   // fractions
};
__device__ __forceinline__ Tap tap(const RenderArgs &r, V3 uvw) {
    const float u[3] = {uvw.x * static_cast<float>(r.n[0]) - 0.5f, uvw.y * static_cast<float>(r.n[1]) - 0.5f,
                        uvw.z * static_cast<float>(r.n[2]) - 0.5f};
    int lo[3], hi[3];
    Tap t;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float fl = floorf(u[a]);
        t.f[a] = u[a] - fl;
        // ray points lie within a step of the clip box; the clamp keeps the conversion defined
        // for any input (indices below -1 or past n - 1 select the edge texel either way)
        const int i = static_cast<int>(fminf(fmaxf(fl, -1.0f), static_cast<float>(r.n[a])));
        lo[a] = min(max(i, 0), r.n[a] - 1);
        hi[a] = min(max(i + 1, 0), r.n[a] - 1);
    }
    const size_t sy = static_cast<size_t>(r.n[0]), sz = sy * static_cast<size_t>(r.n[1]);
#pragma unroll
    for (int c = 0; c < 8; ++c)
        t.p[c] = static_cast<size_t>((c & 4) ? hi[2] : lo[2]) * sz + static_cast<size_t>((c & 2) ? hi[1] : lo[1]) * sy +
                 static_cast<size_t>((c & 1) ? hi[0] : lo[0]);
    return t;
}
// x, then y, then z
__device__ __forceinline__ float blend8(const Tap &t, const float (&v)[8]) {
    const float x00 = mixf(v[0], v[1], t.f[0]), x10 = mixf(v[2], v[3], t.f[0]);
    const float x01 = mixf(v[4], v[5], t.f[0]), x11 = mixf(v[6], v[7], t.f[0]);
    return mixf(mixf(x00, x10, t.f[1]), mixf(x01, x11, t.f[1]), t.f[2]);
}
__device__ __forceinline__ float sample_field(const FieldPtrs &f, int comp, const Tap &t) {
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = field_load(f, comp, t.p[c]);
    return blend8(t, v);
}
__device__ __forceinline__ float sample_plane(const float *__restrict__ plane, const Tap &t) {
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = plane[t.p[c]];
    return blend8(t, v);
}

// slab test, (enter, exit); enter > exit is a miss (raymarch.frag:38-47)
__device__ __forceinline__ void box_hit(V3 ro, V3 rd, V3 bmin, V3 bmax, float &t0, float &t1) {
    auto safe = [](float d) { return fabsf(d) < 1e-8f ? 1e-8f : d; };
    const V3 inv = v3(1.0f / safe(rd.x), 1.0f / safe(rd.y), 1.0f / safe(rd.z));
    const V3 a = mul(bmin - ro, inv), b = mul(bmax - ro, inv);
    const V3 mn = v3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z));
    const V3 mx = v3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z));
    t0 = fmaxf(fmaxf(mn.x, mn.y), mn.z);
    t1 = fminf(fminf(mx.x, mx.y), mx.z);
}

__device__ __forceinline__ V3 shade(V3 n, V3 rd, V3 albedo, float specular, V3 key) {
    const V3 v = rd * -1.0f;
    const float ndl = fmaxf(dot3(n, key), 0.0f);
    const float head = fmaxf(dot3(n, v), 0.0f);
    const V3 hv = unit(key + v);
    const float spec = powf(fmaxf(dot3(n, hv), 0.0f), 48.0f) * specular;
    const float rim = powf(1.0f - head, 3.0f);
    const float k = (0.18f + 0.55f * ndl) + 0.30f * head;
    return (albedo * k + v3(spec, spec, spec)) + v3(0.35f, 0.45f, 0.6f) * rim;
}

__global__ void __launch_bounds__(256) k_raymarch(RenderArgs r, FieldPtrs f, const float *__restrict__ sigma_cell,
                                                  const float *__restrict__ eps_cell, void *out, int rgba8) {
    // 16 x 16 pixel tile per workgroup, 8 x 8 per wavefront
    const int lane = static_cast<int>(threadIdx.x) & 63, w = static_cast<int>(threadIdx.x) >> 6;
    const int px = static_cast<int>(blockIdx.x) * 16 + (w & 1) * 8 + (lane & 7);
    const int py = static_cast<int>(blockIdx.y) * 16 + (w >> 1) * 8 + (lane >> 3);
    if (px >= r.width || py >= r.height) return;
    const float fx = static_cast<float>(px) + 0.5f, fy = static_cast<float>(py) + 0.5f;  // gl_FragCoord
    const float ndc[4] = {2.0f * fx / static_cast<float>(r.width) - 1.0f, 2.0f * fy / static_cast<float>(r.height) - 1.0f,
                          1.0f, 1.0f};
    float far[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        far[i] = ((r.inv_vp[i] * ndc[0] + r.inv_vp[4 + i] * ndc[1]) + r.inv_vp[8 + i] * ndc[2]) + r.inv_vp[12 + i] * ndc[3];
    const V3 ro = v3(r.cam[0], r.cam[1], r.cam[2]);
    const V3 rd = unit(v3(far[0] / far[3], far[1] / far[3], far[2] / far[3]) - ro);

    // background: dark blue gradient over rd.z
    const float h = clamp01(rd.z * 0.5f + 0.5f);
    const V3 bg = mix3(v3(0.004f, 0.006f, 0.014f), v3(0.018f, 0.028f, 0.060f), h);
    float t0, t1;
    box_hit(ro, rd, v3(r.clip_min[0], r.clip_min[1], r.clip_min[2]), v3(r.clip_max[0], r.clip_max[1], r.clip_max[2]),
            t0, t1);
    t0 = fmaxf(t0, 0.0f);
    V3 accum = v3(0.0f, 0.0f, 0.0f);
    float alpha = 0.0f;
    if (t0 < t1) {
        const float voxel = r.vol_size[0] / static_cast<float>(r.n[0]);
        const float ds = fmaxf(r.step_voxels, 0.05f) * voxel;
        const float svox = ds / voxel;
        float t = t0 + ds * gradient_noise(fx, fy);
        // at most 4096 steps (MAX_STEPS); the float clamp bounds the conversion for any input
        const int steps = static_cast<int>(fminf(fmaxf((t1 - t) / ds, 0.0f), 4095.0f)) + 1;
        const V3 vmin = v3(r.vol_min[0], r.vol_min[1], r.vol_min[2]);
        const V3 vsz = v3(r.vol_size[0], r.vol_size[1], r.vol_size[2]);
        const V3 key = unit(v3(0.45f, -0.35f, 0.82f));
        const V3 hstep = v3(1.0f / static_cast<float>(r.n[0]), 1.0f / static_cast<float>(r.n[1]),
                            1.0f / static_cast<float>(r.n[2]));
        for (int i = 0; i < steps; ++i, t += ds) {
            const V3 q = (ro + rd * t) - vmin;
            const V3 uvw = v3(q.x / vsz.x, q.y / vsz.y, q.z / vsz.z);
            const Tap tp = tap(r, uvw);
            const V3 e = v3(sample_field(f, 0, tp), sample_field(f, 1, tp), sample_field(f, 2, tp));
            const V3 b = v3(sample_field(f, 3, tp), sample_field(f, 4, tp), sample_field(f, 5, tp));
            const float sig = sample_plane(sigma_cell, tp), eps = sample_plane(eps_cell, tp);

            // field emission (raymarch.frag:132-165)
            V3 em = v3(0.0f, 0.0f, 0.0f);
            float ext = 0.0f;
            auto glow = [&](float iv, V3 lo, V3 hi, V3 top) {
                V3 col = mix3(lo, hi, iv);
                col = mix3(col, top, smooth(0.8f, 1.0f, iv));
                em = em + col * (iv * iv * 2.0f);
                ext += iv;
            };
            if (r.mode == 0 || r.mode == 1)
                glow(response(len3(e) * r.gain_e), v3(0.02f, 0.22f, 1.0f), v3(0.05f, 0.95f, 1.0f), v3(0.8f, 1.0f, 1.0f));
            if (r.mode == 0 || r.mode == 2)
                glow(response(len3(b) * r.gain_b), v3(1.0f, 0.03f, 0.10f), v3(1.0f, 0.10f, 0.62f), v3(1.0f, 0.85f, 1.0f));
            if (r.mode == 3)
                glow(response(len3(cross3(e, b)) * r.gain_e * r.gain_b), v3(1.0f, 0.35f, 0.02f), v3(1.0f, 0.85f, 0.25f),
                     v3(1.0f, 1.0f, 0.85f));
            if (r.mode == 4) {
                const float v = (r.signed_axis == 0 ? e.x : r.signed_axis == 1 ? e.y : e.z) * r.gain_e;
                const float iv = response(v);
                const V3 col = v >= 0.0f ? mix3(v3(0.02f, 0.3f, 1.0f), v3(0.1f, 1.0f, 1.0f), iv)
                                         : mix3(v3(0.9f, 0.15f, 0.02f), v3(1.0f, 0.65f, 0.1f), iv);
                em = em + col * (iv * iv * 2.0f);
                ext += iv;
            }
            em = em * (0.045f * r.emission);
            ext *= 0.022f * r.density;

            // materials (raymarch.frag:167-185): metal from sigma, glass from eps_r, lossy media
            const float metal = smooth(1.0f, 20.0f, sig);
            const float glass = clamp01((eps - 1.0f) * 0.5f);
            const float lossy = smooth(0.02f, 1.0f, sig) * (1.0f - metal);
            V3 surf = v3(0.0f, 0.0f, 0.0f);
            float mext = 0.0f;
            if ((metal + glass) + lossy > 0.01f) {
                const bool use_sigma = metal >= glass;
                const float *plane = use_sigma ? sigma_cell : eps_cell;
                const float gx = sample_plane(plane, tap(r, uvw + v3(hstep.x, 0.0f, 0.0f))) -
                                 sample_plane(plane, tap(r, uvw - v3(hstep.x, 0.0f, 0.0f)));
                const float gy = sample_plane(plane, tap(r, uvw + v3(0.0f, hstep.y, 0.0f))) -
                                 sample_plane(plane, tap(r, uvw - v3(0.0f, hstep.y, 0.0f)));
                const float gz = sample_plane(plane, tap(r, uvw + v3(0.0f, 0.0f, hstep.z))) -
                                 sample_plane(plane, tap(r, uvw - v3(0.0f, 0.0f, hstep.z)));
                const V3 g = v3(gx, gy, gz);
                const float gl = len3(g);
                // a homogeneous block has no gradient: lit as if facing the viewer
                const V3 n = gl > 1e-5f ? v3(-g.x / gl, -g.y / gl, -g.z / gl) : rd * -1.0f;
                const V3 cm = shade(n, rd, v3(0.50f, 0.53f, 0.58f), 0.9f, key);
                const V3 cg = shade(n, rd, v3(0.30f, 0.65f, 0.70f), 0.6f, key);
                const V3 ca = shade(n, rd, v3(0.35f, 0.22f, 0.12f), 0.1f, key);
                const float wm = metal * r.material_opacity;
                const float wg = glass * r.material_opacity * 0.02f;
                const float wa = lossy * r.material_opacity * 0.4f;
                mext = (wm + wg) + wa;
                const V3 sum = (cm * wm + cg * wg) + ca * wa;
                const float d = fmaxf(mext, 1e-6f);
                surf = v3(sum.x / d, sum.y / d, sum.z / d);
            }

            // front-to-back compositing
            const float af = 1.0f - expf(-ext * svox);
            const float am = 1.0f - expf(-mext * svox);
            const float tr = 1.0f - alpha;
            accum = accum + ((em * svox) * (1.0f - am) + surf * am) * tr;
            alpha += tr * (1.0f - (1.0f - af) * (1.0f - am));
            if (alpha > 0.995f) break;
        }
    }
    // tone curve, display gamma, dither (raymarch.frag:96-101)
    V3 c = accum + bg * (1.0f - alpha);
    const float dither = (gradient_noise(fx + 17.0f, fy + 17.0f) - 0.5f) / 255.0f;
    float o[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        o[k] = 1.0f - expf(-o[k] * r.exposure);
        o[k] = powf(o[k], 1.0f / 2.2f);
        o[k] += dither;
    }
    const size_t pix = static_cast<size_t>(py) * r.width + px;
    if (rgba8) {
        uchar4 u;
        u.x = static_cast<unsigned char>(rintf(clamp01(o[0]) * 255.0f));
        u.y = static_cast<unsigned char>(rintf(clamp01(o[1]) * 255.0f));
        u.z = static_cast<unsigned char>(rintf(clamp01(o[2]) * 255.0f));
        u.w = 255;
        static_cast<uchar4 *>(out)[pix] = u;
    } else {
        static_cast<float4 *>(out)[pix] = make_float4(o[0], o[1], o[2], 1.0f);
    }
}

}  // namespace

hipError_t launch_raymarch(const RenderArgs &r, const FieldPtrs &f, const float *sigma_cell, const float *eps_cell,
                           void *out, int rgba8, hipStream_t s) {
    const dim3 grid((r.width + 15) / 16, (r.height + 15) / 16);
    k_raymarch<<<grid, 256, 0, s>>>(r, f, sigma_cell, eps_cell, out, rgba8);
    return hipGetLastError();
}

}  // namespace emfd
