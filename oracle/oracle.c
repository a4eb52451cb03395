/*
 * CPU oracle for the EM_Lab FDTD hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A scalar C restatement of the reference's compute shaders and of the host code around them,
 * used as the checker for the HIP implementation. Only tests, the smoke check and the
 * benchmark's CPU-baseline leg load this library; the product never does.
 *
 * Every function cites the reference file:line it restates. Floating-point expressions keep
 * the reference's operand order; the file is compiled with -ffp-contract=off so that no
 * multiply-add is fused behind the reader's back.
 *
 * Pinning status (what checks each function; the tests live in tests/test_oracle_*.py,
 * tests/test_storage_cpu.py and tests/test_coax_cpu.py):
 *   f16 packing          known answers (binary16 RNE) + numpy's RNE conversion on 69k values
 *   configure            clamp table from src/SimulationGrid.cpp:149-156
 *   upload_materials     linear media: every edge average equals the medium at the edge's
 *                        midpoint; one isolated cell touches exactly the 12 edges its geometry
 *                        implies (index-0 clamping at the lower faces)
 *   update_B / update_E  exact discrete plane wave, discrete energy invariant, lossy decay
 *                        closed form, phi continuity across its branch point
 *   apply_sources        sum J = J0 x Poisson-summed footprint, cell-list brute force
 *   signal               direct evaluation of the formula at sample times
 *   port / ntff EMA      two-stage EMA closed form for a pure sinusoid
 *   e_at_node / b_at_node / ntff_samples
 *                        linear fields at each component's stagger return the analytic node
 *                        value exactly (a half-cell offset fails)
 *   ntff pattern         Hertzian dipole closed-form near field -> D = 1.5 sin^2, P = k^2/12pi;
 *                        f32 order vs f64 brute force
 *   flux / probe         uniform and linear fields, closed-form plane and box flux
 *   sweep                resistor closed form (Z = R), recurrence vs direct polar
 *   line_accumulate      Laplace / gradient / Stokes identities of the term lists
 *   step_storage         numpy float16 / torch bfloat16 conversions, incl. ties and subnormals
 *   cpml_*               closed-form reflection of a graded stretched-coordinate layer, held to
 *                        the lattice's sampling excess (10-20% in ln R at 8 cells); parity
 *                        unpinned beyond that bound
 * A planted-mistake run (tests/test_oracle_mutations.py) shows each pin rejecting the mistakes
 * it is meant to catch.
 * GLSL built-ins are driver-defined in the reference. exp is or_exp below (the library
 * evaluates the same method); sin and cos in the pattern are the C library's, so that
 * boundary is parity unpinned at the ulp level (patterns are compared in dB).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(g, x, y, z) ((((size_t)(z)) * (size_t)(g)->ny + (size_t)(y)) * (size_t)(g)->nx + (size_t)(x))

static inline float fmaxf_(float a, float b) { return a > b ? a : b; }

/* The GLSL built-in exp is driver-defined (no version pin in the reference). Both this oracle
 * and the GPU library evaluate it with the same published method so that their results can be
 * compared bit for bit: Cody-Waite reduction x = n ln2 + r (ln2 as a 16-bit head plus a tail),
 * Taylor series of e^r to degree 7 by Horner's rule, exact scaling by 2^n. Within 2 ulp. */
float or_exp(float x) {
    if (x != x) return x;
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
    return ldexpf(p, (int)n);
}
static inline float fminf_(float a, float b) { return a < b ? a : b; }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ------------------------------------------------------------------------------------------
 * binary16 conversion, round to nearest even (glm::packHalf2x16 as used at
 * src/SimulationGrid.cpp:190-191,260-261,276,287-288)
 * ------------------------------------------------------------------------------------------ */
uint16_t or_f32_to_f16(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t a = x & 0x7fffffffu;
    if (a >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (a > 0x7f800000u ? 0x200u : 0u));
    if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* >= 65520 rounds to infinity */
    if (a >= 0x38800000u) {                                  /* normal half */
        uint32_t e = (a >> 23) - 127u + 15u;
        uint32_t m = a & 0x7fffffu;
        uint32_t h = (e << 10) | (m >> 13);
        uint32_t rem = m & 0x1fffu;
        if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    if (a < 0x33000000u) return (uint16_t)sign; /* below 2^-25 (ties at 2^-25 go to even zero) */
    {
        uint32_t e = a >> 23;
        uint32_t m = (a & 0x7fffffu) | 0x800000u;
        uint32_t s = 126u - e; /* value / 2^-24 = m >> s */
        uint32_t q = m >> s;
        uint32_t rem = m & ((1u << s) - 1u);
        uint32_t half = 1u << (s - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(sign | q);
    }
}

float or_f16_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1fu;
    uint32_t m = (uint32_t)h & 0x3ffu;
    uint32_t out;
    if (e == 0) {
        if (m == 0) {
            out = sign;
        } else { /* subnormal: renormalise */
            int k = -1;
            do { ++k; m <<= 1; } while (!(m & 0x400u));
            out = sign | ((uint32_t)(127 - 15 - k) << 23) | ((m & 0x3ffu) << 13);
        }
    } else if (e == 31) {
        out = sign | 0x7f800000u | (m << 13);
    } else {
        out = sign | ((e - 15u + 127u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &out, 4);
    return f;
}

uint32_t or_pack_half2(float lo, float hi) {
    return (uint32_t)or_f32_to_f16(lo) | ((uint32_t)or_f32_to_f16(hi) << 16);
}

static inline void unpack_half2(uint32_t u, float *lo, float *hi) {
    *lo = or_f16_to_f32((uint16_t)(u & 0xffffu));
    *hi = or_f16_to_f32((uint16_t)(u >> 16));
}

/* ------------------------------------------------------------------------------------------
 * SimulationGrid::configure clamps (src/SimulationGrid.cpp:24,149-156)
 * ------------------------------------------------------------------------------------------ */
void or_configure(int nx, int ny, int nz, float courant, int pml, float sigma_max, float grading, or_grid *out) {
    const float kMaxCourant = 0.99f / 1.7320508f;
    pml = clampi(pml, 0, 64);
    int min_size = 2 * pml + 8;
    out->nx = nx > min_size ? nx : min_size;
    out->ny = ny > min_size ? ny : min_size;
    out->nz = nz > min_size ? nz : min_size;
    out->dt = courant < 0.05f ? 0.05f : (courant > kMaxCourant ? kMaxCourant : courant);
    out->pml = pml;
    out->sigma_max = sigma_max > 0.0f ? sigma_max : 0.0f;
    out->grading = grading < 0.25f ? 0.25f : (grading > 12.0f ? 12.0f : grading);
}

/* ------------------------------------------------------------------------------------------
 * common.glsl (shaders/common.glsl:23-54)
 * ------------------------------------------------------------------------------------------ */
float or_pml_profile(const or_grid *g, float depth) {
    if (depth <= 0.0f) return 0.0f;
    float u = fminf_(depth / (float)g->pml, 1.0f); /* depth/0 = inf -> 1 */
    return g->sigma_max * (or_exp(g->grading * u) - 1.0f) / (or_exp(g->grading) - 1.0f);
}

float or_pml_sigma(const or_grid *g, float x, float y, float z) {
    const float L = (float)g->pml;
    const float hx = (float)(g->nx - 1) - L, hy = (float)(g->ny - 1) - L, hz = (float)(g->nz - 1) - L;
    const float dx = fmaxf_(L - x, x - hx), dy = fmaxf_(L - y, y - hy), dz = fmaxf_(L - z, z - hz);
    return or_pml_profile(g, dx) + or_pml_profile(g, dy) + or_pml_profile(g, dz);
}

int or_in_pml(const or_grid *g, int x, int y, int z) {
    const int L = g->pml;
    return x < L || y < L || z < L || x >= g->nx - (L + 1) || y >= g->ny - (L + 1) || z >= g->nz - (L + 1);
}

float or_phi(float x) {
    const float series = 1.0f - x * (0.5f - x * (1.0f / 6.0f - x * (1.0f / 24.0f)));
    const float exact = (1.0f - or_exp(-x)) / fmaxf_(x, 1e-12f);
    return x >= 1e-2f ? exact : series; /* mix(series, exact, step(1e-2, x)) */
}

/* ------------------------------------------------------------------------------------------
 * uploadMaterials (src/SimulationGrid.cpp:209-295)
 * ------------------------------------------------------------------------------------------ */
int or_upload_materials(const or_grid *g, const float *sig, const float *eps, const or_edge_override *edges,
                        size_t n_edges, const or_wire_coef *wires, size_t n_wire, float *E, float *B,
                        uint32_t *edge_mat, uint32_t *wire_tex) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t i = 0; i < n; ++i) {
        E[4 * i + 0] = E[4 * i + 1] = E[4 * i + 2] = 0.0f;
        E[4 * i + 3] = sig[i];
        B[4 * i + 0] = B[4 * i + 1] = B[4 * i + 2] = 0.0f;
        B[4 * i + 3] = eps[i];
    }
    float *es = (float *)malloc(n * 3 * sizeof(float));
    float *ee = (float *)malloc(n * 3 * sizeof(float));
#define CIDX(x, y, z) IDX(g, (x) > 0 ? (x) : 0, (y) > 0 ? (y) : 0, (z) > 0 ? (z) : 0)
#define AVG(v, a, b, c, d) (0.25f * ((v)[a] + (v)[b] + (v)[c] + (v)[d]))
    for (int z = 0; z < g->nz; ++z)
        for (int y = 0; y < g->ny; ++y)
            for (int x = 0; x < g->nx; ++x) {
                const size_t c00 = IDX(g, x, y, z);
                const size_t ex[4] = {c00, CIDX(x, y - 1, z), CIDX(x, y, z - 1), CIDX(x, y - 1, z - 1)};
                const size_t ey[4] = {c00, CIDX(x - 1, y, z), CIDX(x, y, z - 1), CIDX(x - 1, y, z - 1)};
                const size_t ez[4] = {c00, CIDX(x - 1, y, z), CIDX(x, y - 1, z), CIDX(x - 1, y - 1, z)};
                es[3 * c00 + 0] = AVG(sig, ex[0], ex[1], ex[2], ex[3]);
                es[3 * c00 + 1] = AVG(sig, ey[0], ey[1], ey[2], ey[3]);
                es[3 * c00 + 2] = AVG(sig, ez[0], ez[1], ez[2], ez[3]);
                ee[3 * c00 + 0] = AVG(eps, ex[0], ex[1], ex[2], ex[3]);
                ee[3 * c00 + 1] = AVG(eps, ey[0], ey[1], ey[2], ey[3]);
                ee[3 * c00 + 2] = AVG(eps, ez[0], ez[1], ez[2], ez[3]);
            }
#undef CIDX
#undef AVG
    for (size_t k = 0; k < n_edges; ++k) {
        const or_edge_override *o = &edges[k];
        if (o->texel[0] < 0 || o->texel[1] < 0 || o->texel[2] < 0 || o->texel[0] >= g->nx || o->texel[1] >= g->ny ||
            o->texel[2] >= g->nz || o->axis < 0 || o->axis > 2)
            continue;
        es[3 * IDX(g, o->texel[0], o->texel[1], o->texel[2]) + (size_t)o->axis] = fmaxf_(o->sigma, 0.0f);
    }
    for (size_t i = 0; i < n; ++i) {
        const float sx = fminf_(es[3 * i + 0], 6.0e4f), sy = fminf_(es[3 * i + 1], 6.0e4f),
                    sz = fminf_(es[3 * i + 2], 6.0e4f);
        edge_mat[4 * i + 0] = or_pack_half2(sx, sy);
        edge_mat[4 * i + 1] = or_pack_half2(sz, ee[3 * i + 0]);
        edge_mat[4 * i + 2] = or_pack_half2(ee[3 * i + 1], ee[3 * i + 2]);
        edge_mat[4 * i + 3] = 0u;
    }
    free(es);
    free(ee);

    const int thin = n_wire > 0;
    const uint32_t ones = or_pack_half2(1.0f, 1.0f);
    if (!thin) {
        if (wire_tex) { wire_tex[0] = wire_tex[1] = wire_tex[2] = ones; wire_tex[3] = 0u; }
        return 0;
    }
    float *f = (float *)malloc(n * 6 * sizeof(float));
    for (size_t i = 0; i < n * 6; ++i) f[i] = 1.0f;
    for (size_t k = 0; k < n_wire; ++k) {
        const or_wire_coef *w = &wires[k];
        if (w->texel[0] < 0 || w->texel[1] < 0 || w->texel[2] < 0 || w->texel[0] >= g->nx || w->texel[1] >= g->ny ||
            w->texel[2] >= g->nz || w->component < 0 || w->component > 2)
            continue;
        f[6 * IDX(g, w->texel[0], w->texel[1], w->texel[2]) + 2 * (size_t)w->component + (w->term == 0 ? 0 : 1)] =
            w->factor;
    }
    for (size_t i = 0; i < n; ++i) {
        wire_tex[4 * i + 0] = or_pack_half2(f[6 * i + 0], f[6 * i + 1]);
        wire_tex[4 * i + 1] = or_pack_half2(f[6 * i + 2], f[6 * i + 3]);
        wire_tex[4 * i + 2] = or_pack_half2(f[6 * i + 4], f[6 * i + 5]);
        wire_tex[4 * i + 3] = 0u;
    }
    free(f);
    return 1;
}

/* ------------------------------------------------------------------------------------------
 * Sources: SourceDesc::signal (src/SimulationGrid.cpp:99-130) and setSources (:297-357)
 * ------------------------------------------------------------------------------------------ */
double or_signal(const or_source_desc *s, double tAbs) {
    const double wl = (double)s->wavelength > 1e-3 ? (double)s->wavelength : 1e-3;
    const double delay = (double)s->delay_periods > 0.0 ? (double)s->delay_periods : 0.0;
    const double t = tAbs - delay * wl;
    if (!s->enabled || t < 0.0) return 0.0;
    const double period = wl;
    const double w = 6.283185307179586476925286766559 / wl;
    if (s->waveform == 0) {
        const double rampT = ((double)s->ramp_periods > 0.0 ? (double)s->ramp_periods : 0.0) * period;
        double ramp = 1.0;
        if (rampT > 0.0 && t < rampT) {
            const double r = t / rampT;
            ramp = r * r * r * (10.0 + r * (-15.0 + 6.0 * r));
        }
        return ramp * sin(w * t + (double)s->phase);
    }
    const double tau = ((double)s->ramp_periods > 0.1 ? (double)s->ramp_periods : 0.1) * period;
    const double t0 = 3.0 * tau;
    double tl = t;
    if (s->repeat_periods > 0.0f) {
        double rep = (double)s->repeat_periods * period;
        if (rep < 2.0 * t0) rep = 2.0 * t0;
        tl = fmod(t, rep);
    }
    const double u = (tl - t0) / tau;
    if (fabs(u) > 6.0) return 0.0;
    return exp(-u * u) * sin(w * (tl - t0) + (double)s->phase);
}

typedef struct { uint32_t cell, src; } pair_t;
static int cmp_pair(const void *a, const void *b) {
    const pair_t *p = (const pair_t *)a, *q = (const pair_t *)b;
    if (p->cell != q->cell) return p->cell < q->cell ? -1 : 1;
    if (p->src != q->src) return p->src < q->src ? -1 : 1;
    return 0;
}

int or_set_sources(const or_grid *g, const or_source_desc *src, int n_src, float *packed, int *kept_index, int *cells,
                   int *n_cells, int *lists, int *n_list) {
    float pk[256 * 16];
    int kept = 0;
    for (int i = 0; i < n_src && kept < 256; ++i) {
        const or_source_desc *s = &src[i];
        const float len = sqrtf(s->polarization[0] * s->polarization[0] + s->polarization[1] * s->polarization[1] +
                                s->polarization[2] * s->polarization[2]);
        if (!s->enabled || !(len > 1e-6f) || !isfinite(len)) continue;
        float *p = &pk[16 * kept];
        for (int a = 0; a < 3; ++a) {
            p[0 + a] = s->center[a];
            float h = s->half_extent[a];
            p[4 + a] = h < 0.0f ? 0.0f : (h > 1.0e5f ? 1.0e5f : h);
            p[8 + a] = s->polarization[a] / len;
            p[12 + a] = fmaxf_(s->softness[a], 0.25f);
        }
        p[3] = 0.0f;
        p[7] = 0.0f;
        p[11] = s->amplitude;
        p[15] = 0.0f;
        if (kept_index) kept_index[kept] = i;
        ++kept;
    }
    if (packed) memcpy(packed, pk, sizeof(float) * 16 * (size_t)kept);

    size_t cap = 0;
    for (int i = 0; i < kept; ++i) {
        const float *p = &pk[16 * i];
        size_t vol = 1;
        for (int a = 0; a < 3; ++a) {
            const int na = a == 0 ? g->nx : (a == 1 ? g->ny : g->nz);
            int lo = clampi((int)floorf(p[a] - p[4 + a] - 3.5f * p[12 + a]) - 1, 0, na - 1);
            int hi = clampi((int)ceilf(p[a] + p[4 + a] + 3.5f * p[12 + a]), 0, na - 1);
            vol *= (size_t)(hi >= lo ? hi - lo + 1 : 0);
        }
        cap += vol;
    }
    pair_t *pairs = (pair_t *)malloc((cap ? cap : 1) * sizeof(pair_t));
    size_t np = 0;
    for (int i = 0; i < kept; ++i) {
        const float *p = &pk[16 * i];
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            const int na = a == 0 ? g->nx : (a == 1 ? g->ny : g->nz);
            lo[a] = clampi((int)floorf(p[a] - p[4 + a] - 3.5f * p[12 + a]) - 1, 0, na - 1);
            hi[a] = clampi((int)ceilf(p[a] + p[4 + a] + 3.5f * p[12 + a]), 0, na - 1);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x) {
                    pairs[np].cell = (uint32_t)((z * g->ny + y) * g->nx + x);
                    pairs[np].src = (uint32_t)i;
                    ++np;
                }
    }
    qsort(pairs, np, sizeof(pair_t), cmp_pair);
    int nc = 0, nl = 0;
    for (size_t k = 0; k < np; ++k) {
        if (k == 0 || pairs[k].cell != pairs[k - 1].cell) {
            const uint32_t idx = pairs[k].cell;
            if (cells) {
                cells[4 * nc + 0] = (int)(idx % (uint32_t)g->nx);
                cells[4 * nc + 1] = (int)((idx / (uint32_t)g->nx) % (uint32_t)g->ny);
                cells[4 * nc + 2] = (int)(idx / (uint32_t)(g->nx * g->ny));
                cells[4 * nc + 3] = nl;
            }
            ++nc;
        }
        if (lists) lists[nl] = (int)pairs[k].src;
        ++nl;
    }
    if (cells) {
        cells[4 * nc + 0] = cells[4 * nc + 1] = cells[4 * nc + 2] = 0;
        cells[4 * nc + 3] = nl;
    }
    free(pairs);
    *n_cells = nc;
    *n_list = nl;
    return kept;
}

/* ------------------------------------------------------------------------------------------
 * update_B.comp (shaders/update_B.comp:31-70)
 * ------------------------------------------------------------------------------------------ */
void or_update_B(const or_grid *g, float *B, const float *E, const uint32_t *wire_tex, int thin_wires) {
    const float dt = g->dt;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->nz; ++z)
        for (int y = 0; y < g->ny; ++y)
            for (int x = 0; x < g->nx; ++x) {
                const size_t p = IDX(g, x, y, z);
                const float *e0 = &E[4 * p];
                const float zero[3] = {0.0f, 0.0f, 0.0f};
                const float *ex = x + 1 < g->nx ? &E[4 * IDX(g, x + 1, y, z)] : zero;
                const float *ey = y + 1 < g->ny ? &E[4 * IDX(g, x, y + 1, z)] : zero;
                const float *ez = z + 1 < g->nz ? &E[4 * IDX(g, x, y, z + 1)] : zero;
                float t0[3] = {ey[2] - e0[2], ez[0] - e0[0], ex[1] - e0[1]};
                float t1[3] = {ez[1] - e0[1], ex[2] - e0[2], ey[0] - e0[0]};
                if (thin_wires) {
                    for (int c = 0; c < 3; ++c) {
                        float w0, w1;
                        unpack_half2(wire_tex[4 * p + (size_t)c], &w0, &w1);
                        t0[c] *= w0;
                        t1[c] *= w1;
                    }
                }
                float curl[3] = {t0[0] - t1[0], t0[1] - t1[1], t0[2] - t1[2]};
                float *b = &B[4 * p];
                if (or_in_pml(g, x, y, z)) {
                    const float px = (float)x, py = (float)y, pz = (float)z;
                    const float sm[3] = {or_pml_sigma(g, px + 0.0f, py + 0.5f, pz + 0.5f),
                                         or_pml_sigma(g, px + 0.5f, py + 0.0f, pz + 0.5f),
                                         or_pml_sigma(g, px + 0.5f, py + 0.5f, pz + 0.0f)};
                    for (int c = 0; c < 3; ++c) {
                        const float xx = sm[c] * dt;
                        b[c] = or_exp(-xx) * b[c] - (dt * or_phi(xx)) * curl[c];
                    }
                } else {
                    for (int c = 0; c < 3; ++c) b[c] = b[c] - dt * curl[c];
                }
            }
}

/* edge coefficients as update_E.comp:46-61 / apply_sources.comp:64-75 */
static void edge_sigma_eps(const or_grid *g, const uint32_t *edge_mat, int x, int y, int z, float sigma[3],
                           float epsr[3]) {
    const size_t p = IDX(g, x, y, z);
    float m0x, m0y, m1x, m1y, m2x, m2y;
    unpack_half2(edge_mat[4 * p + 0], &m0x, &m0y);
    unpack_half2(edge_mat[4 * p + 1], &m1x, &m1y);
    unpack_half2(edge_mat[4 * p + 2], &m2x, &m2y);
    sigma[0] = fmaxf_(m0x, 0.0f);
    sigma[1] = fmaxf_(m0y, 0.0f);
    sigma[2] = fmaxf_(m1x, 0.0f);
    epsr[0] = fmaxf_(m1y, 1.0f);
    epsr[1] = fmaxf_(m2x, 1.0f);
    epsr[2] = fmaxf_(m2y, 1.0f);
    if (or_in_pml(g, x, y, z)) {
        const float px = (float)x, py = (float)y, pz = (float)z;
        sigma[0] += or_pml_sigma(g, px + 0.5f, py + 0.0f, pz + 0.0f);
        sigma[1] += or_pml_sigma(g, px + 0.0f, py + 0.5f, pz + 0.0f);
        sigma[2] += or_pml_sigma(g, px + 0.0f, py + 0.0f, pz + 0.5f);
    }
}

/* ------------------------------------------------------------------------------------------
 * update_E.comp (shaders/update_E.comp:31-66)
 * ------------------------------------------------------------------------------------------ */
void or_update_E(const or_grid *g, float *E, const float *B, const uint32_t *edge_mat) {
    const float dt = g->dt;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < g->nz; ++z)
        for (int y = 0; y < g->ny; ++y)
            for (int x = 0; x < g->nx; ++x) {
                const size_t p = IDX(g, x, y, z);
                const float zero[3] = {0.0f, 0.0f, 0.0f};
                const float *b0 = &B[4 * p];
                const float *bx = x > 0 ? &B[4 * IDX(g, x - 1, y, z)] : zero;
                const float *by = y > 0 ? &B[4 * IDX(g, x, y - 1, z)] : zero;
                const float *bz = z > 0 ? &B[4 * IDX(g, x, y, z - 1)] : zero;
                const float curl[3] = {(b0[2] - by[2]) - (b0[1] - bz[1]), (b0[0] - bz[0]) - (b0[2] - bx[2]),
                                       (b0[1] - bx[1]) - (b0[0] - by[0])};
                float sigma[3], epsr[3];
                edge_sigma_eps(g, edge_mat, x, y, z, sigma, epsr);
                float *e = &E[4 * p];
                for (int c = 0; c < 3; ++c) {
                    const float dtEps = dt / epsr[c];
                    const float xx = sigma[c] * dtEps;
                    const float ca = or_exp(-xx);
                    const float cb = dtEps * or_phi(xx);
                    e[c] = ca * e[c] + cb * curl[c];
                }
            }
}

/* ------------------------------------------------------------------------------------------
 * apply_sources.comp (shaders/apply_sources.comp:37-79)
 * ------------------------------------------------------------------------------------------ */
static float footprint(const float *s, float x, float y, float z) {
    const float q[3] = {x, y, z};
    float d[3];
    for (int a = 0; a < 3; ++a) d[a] = fmaxf_(fabsf(q[a] - s[a]) - s[4 + a], 0.0f) / s[12 + a];
    return or_exp(-(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
}

void or_apply_sources(const or_grid *g, float *E, const uint32_t *edge_mat, const float *packed, const float *signals,
                      int signal_base, const int *cells, int n_cells, const int *lists) {
    const float dt = g->dt;
    for (int id = 0; id < n_cells; ++id) {
        const int *entry = &cells[4 * id];
        const int end = cells[4 * (id + 1) + 3];
        const int x = entry[0], y = entry[1], z = entry[2];
        const float cx = (float)x, cy = (float)y, cz = (float)z;
        float J[3] = {0.0f, 0.0f, 0.0f};
        for (int k = entry[3]; k < end; ++k) {
            const int i = lists[k];
            const float signal = signals[signal_base + i];
            if (signal == 0.0f) continue;
            const float *s = &packed[16 * i];
            const float w[3] = {footprint(s, cx + 0.5f, cy + 0.0f, cz + 0.0f), footprint(s, cx + 0.0f, cy + 0.5f, cz + 0.0f),
                                footprint(s, cx + 0.0f, cy + 0.0f, cz + 0.5f)};
            const float as = s[11] * signal;
            for (int c = 0; c < 3; ++c) J[c] += s[8 + c] * as * w[c];
        }
        if (J[0] == 0.0f && J[1] == 0.0f && J[2] == 0.0f) continue;
        float sigma[3], epsr[3];
        edge_sigma_eps(g, edge_mat, x, y, z, sigma, epsr);
        float *e = &E[4 * IDX(g, x, y, z)];
        for (int c = 0; c < 3; ++c) {
            const float dtEps = dt / epsr[c];
            const float cb = dtEps * or_phi(sigma[c] * dtEps);
            e[c] = e[c] - cb * J[c];
        }
    }
}

/* clear_fields.comp (shaders/clear_fields.comp:11-16) */
void or_clear_fields(const or_grid *g, float *E, float *B) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t i = 0; i < n; ++i) {
        E[4 * i + 0] = E[4 * i + 1] = E[4 * i + 2] = 0.0f;
        B[4 * i + 0] = B[4 * i + 1] = B[4 * i + 2] = 0.0f;
    }
}

/* ------------------------------------------------------------------------------------------
 * 16-bit field storage (no reference counterpart): values are kept as binary16 or bfloat16 of
 * value / 2^k, rounded to nearest even, and every update reads them back as float.
 * ------------------------------------------------------------------------------------------ */
float or_round_f16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    const uint32_t sign = u & 0x80000000u, a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return f;                 /* inf, nan */
    uint32_t r;
    if (a >= 0x477ff000u) {                         /* >= 65520: rounds to infinity */
        r = 0x7f800000u;
    } else if (a < 0x38800000u) {                   /* below 2^-14: binary16 subnormal, quantum 2^-24 */
        float v;
        memcpy(&v, &a, 4);
        v = rintf(v * 16777216.0f) * 5.9604644775390625e-8f;
        memcpy(&r, &v, 4);
    } else {                                        /* normal: keep 10 fraction bits */
        r = (a + 0xfffu + ((a >> 13) & 1u)) & ~0x1fffu;
    }
    r |= sign;
    float o;
    memcpy(&o, &r, 4);
    return o;
}

float or_round_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return f;  /* nan */
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    float o;
    memcpy(&o, &u, 4);
    return o;
}

/* Rounds the three components of every texel of F (4 floats per texel) to the storage. */
static void quantize(const or_grid *g, float *F, int storage, const int *k3) {
    if (storage == 0) return;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    float s[3], inv[3];
    for (int c = 0; c < 3; ++c) {
        s[c] = ldexpf(1.0f, k3[c]);
        inv[c] = ldexpf(1.0f, -k3[c]);
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            const float v = F[4 * i + c] * inv[c];
            F[4 * i + c] = (storage == 1 ? or_round_f16(v) : or_round_bf16(v)) * s[c];
        }
}

void or_quantize_fields(const or_grid *g, float *E, float *B, int storage, const int scale_log2[6]) {
    quantize(g, E, storage, scale_log2);
    quantize(g, B, storage, scale_log2 + 3);
}

/* SimulationGrid::step (src/SimulationGrid.cpp:368-433); storage 1/2 rounds every pass's
 * output to binary16 / bfloat16 as the 16-bit storage of the library does. */
void or_step_storage(const or_grid *g, float *E, float *B, const uint32_t *edge_mat, const uint32_t *wire_tex,
                     int thin_wires, const or_source_desc *src, int n_src, int count, uint64_t *step_count,
                     int storage, const int *scale_log2) {
    if (count <= 0) return;
    float packed[256 * 16];
    int kept_index[256];
    int n_cells = 0, n_list = 0;
    int kept = or_set_sources(g, src, n_src, NULL, NULL, NULL, &n_cells, NULL, &n_list);
    int *cells = (int *)malloc(sizeof(int) * 4 * (size_t)(n_cells + 1));
    int *lists = (int *)malloc(sizeof(int) * (size_t)(n_list > 0 ? n_list : 1));
    kept = or_set_sources(g, src, n_src, packed, kept_index, cells, &n_cells, lists, &n_list);
    float signals[256];
    for (int k = 0; k < count; ++k) {
        const double t = ((double)(*step_count) + 0.5) * (double)g->dt;
        for (int s = 0; s < kept; ++s) signals[s] = (float)or_signal(&src[kept_index[s]], t);
        or_update_B(g, B, E, wire_tex, thin_wires);
        if (storage) quantize(g, B, storage, scale_log2 + 3);
        or_update_E(g, E, B, edge_mat);
        if (storage) quantize(g, E, storage, scale_log2);
        if (n_cells > 0) {
            or_apply_sources(g, E, edge_mat, packed, signals, 0, cells, n_cells, lists);
            if (storage) quantize(g, E, storage, scale_log2);
        }
        ++(*step_count);
    }
    free(cells);
    free(lists);
}

void or_step(const or_grid *g, float *E, float *B, const uint32_t *edge_mat, const uint32_t *wire_tex, int thin_wires,
             const or_source_desc *src, int n_src, int count, uint64_t *step_count) {
    or_step_storage(g, E, B, edge_mat, wire_tex, thin_wires, src, n_src, count, step_count, 0, NULL);
}

/* ------------------------------------------------------------------------------------------
 * Convolutional PML (a library extension with no reference counterpart; the reference has the
 * graded absorber only). The textbook recursive-convolution form: J. A. Roden and S. D. Gedney,
 * Microwave Opt. Technol. Lett. 27(5), 334-339 (2000), with the complex-frequency-shifted
 * stretch s = kappa + sigma / (alpha + j w) (M. Kuzuoglu and R. Mittra, IEEE MGWL 6(12), 1996),
 * kappa = 1. In units where eps0 = mu0 = 1 and the cell is 1:
 *   b = exp(-(sigma + alpha) dt),  a = sigma / (sigma + alpha) (b - 1),
 *   psi^{n+1} = b psi^n + a (d_u F),  and the update's d_u F gains + psi.
 * Profiles over the normalised depth d in (0, 1] of the graded absorber's slabs:
 * sigma = sigma_max d^m, alpha = alpha_max (1 - d); 0 (b = 1, a = 0) outside.
 * Each axis u holds psi for the S = 2L + 1 lattice planes i = 0 .. L-1 and n-1-L .. n-1 (the
 * half-integer B positions of plane n-1-L are inside the absorber): 4 components (B_v, B_w, E_v,
 * E_w with v = u+1, w = u+2 mod 3) over the slab's cells, x fastest.
 * ------------------------------------------------------------------------------------------ */
static int cpml_dims(const or_grid *g, int u, int dims[3]) {
    const int n[3] = {g->nx, g->ny, g->nz};
    for (int a = 0; a < 3; ++a) dims[a] = n[a];
    dims[u] = 2 * g->pml + 1;
    return dims[0] * dims[1] * dims[2];
}

static int cpml_plane(const or_grid *g, int u, int s) {
    const int n[3] = {g->nx, g->ny, g->nz};
    return s < g->pml ? s : n[u] - 1 - 2 * g->pml + s;
}

/* coef[((u * 2 + kind) * stride + i) * 2 + (0: b, 1: a)], kind 0 = E position i, 1 = B position
 * i + 1/2; stride = max(nx, ny, nz). */
void or_cpml_coefficients(const or_grid *g, double sigma_max, double grading, double alpha_max, float *coef) {
    const int n[3] = {g->nx, g->ny, g->nz};
    const int stride = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
    const double L = (double)g->pml, dt = (double)g->dt;
    for (int u = 0; u < 3; ++u)
        for (int kind = 0; kind < 2; ++kind)
            for (int i = 0; i < stride; ++i) {
                float *c = &coef[((size_t)(u * 2 + kind) * stride + i) * 2];
                c[0] = 1.0f;
                c[1] = 0.0f;
                if (i >= n[u] || g->pml <= 0) continue;
                const double x = (double)i + (kind ? 0.5 : 0.0);
                const double hi = (double)(n[u] - 1) - L;
                double d = (L - x > x - hi ? L - x : x - hi) / L;
                if (d <= 0.0) continue;
                if (d > 1.0) d = 1.0;
                const double sigma = sigma_max * pow(d, grading), alpha = alpha_max * (1.0 - d);
                const double bb = exp(-(sigma + alpha) * dt);
                c[0] = (float)bb;
                c[1] = sigma + alpha > 0.0 ? (float)(sigma / (sigma + alpha) * (bb - 1.0)) : 0.0f;
            }
}

/* After the Faraday update: B_v += dt psi(d_u E_w), B_w -= dt psi(d_u E_v), forward
 * differences (E beyond the last plane is 0); axis x, then y, then z. */
void or_cpml_b(const or_grid *g, float *B, const float *E, const float *coef, float *psi) {
    const int n[3] = {g->nx, g->ny, g->nz};
    const int stride = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
    const float dt = g->dt;
    size_t off = 0;
    for (int u = 0; u < 3; ++u) {
        const int v = (u + 1) % 3, w = (u + 2) % 3;
        int dims[3];
        const int cells = cpml_dims(g, u, dims);
        float *pv = psi + off, *pw = psi + off + cells;
#pragma omp parallel for schedule(static)
        for (int t = 0; t < cells; ++t) {
            int q[3] = {t % dims[0], (t / dims[0]) % dims[1], t / (dims[0] * dims[1])};
            q[u] = cpml_plane(g, u, q[u]);
            int r[3] = {q[0], q[1], q[2]};
            r[u] += 1;
            const float *e0 = &E[4 * IDX(g, q[0], q[1], q[2])];
            const int in = r[u] < n[u];
            const float ew1 = in ? E[4 * IDX(g, r[0], r[1], r[2]) + w] : 0.0f;
            const float ev1 = in ? E[4 * IDX(g, r[0], r[1], r[2]) + v] : 0.0f;
            const float *c = &coef[((size_t)(u * 2 + 1) * stride + q[u]) * 2];
            pv[t] = c[0] * pv[t] + c[1] * (ew1 - e0[w]);
            pw[t] = c[0] * pw[t] + c[1] * (ev1 - e0[v]);
            float *b = &B[4 * IDX(g, q[0], q[1], q[2])];
            b[v] = b[v] + dt * pv[t];
            b[w] = b[w] - dt * pw[t];
        }
        off += 4 * (size_t)cells;
    }
}

/* After the Ampere update: E_v -= Cb_v psi(d_u B_w), E_w += Cb_w psi(d_u B_v), backward
 * differences (B before the first plane is 0), Cb of the edge as update_E.comp forms it. */
void or_cpml_e(const or_grid *g, float *E, const float *B, const uint32_t *edge_mat, const float *coef, float *psi) {
    const int n[3] = {g->nx, g->ny, g->nz};
    const int stride = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
    const float dt = g->dt;
    size_t off = 0;
    for (int u = 0; u < 3; ++u) {
        const int v = (u + 1) % 3, w = (u + 2) % 3;
        int dims[3];
        const int cells = cpml_dims(g, u, dims);
        float *pv = psi + off + 2 * (size_t)cells, *pw = psi + off + 3 * (size_t)cells;
#pragma omp parallel for schedule(static)
        for (int t = 0; t < cells; ++t) {
            int q[3] = {t % dims[0], (t / dims[0]) % dims[1], t / (dims[0] * dims[1])};
            q[u] = cpml_plane(g, u, q[u]);
            int r[3] = {q[0], q[1], q[2]};
            r[u] -= 1;
            const float *b0 = &B[4 * IDX(g, q[0], q[1], q[2])];
            const int in = r[u] >= 0;
            const float bw1 = in ? B[4 * IDX(g, r[0], r[1], r[2]) + w] : 0.0f;
            const float bv1 = in ? B[4 * IDX(g, r[0], r[1], r[2]) + v] : 0.0f;
            const float *c = &coef[((size_t)(u * 2 + 0) * stride + q[u]) * 2];
            pv[t] = c[0] * pv[t] + c[1] * (b0[w] - bw1);
            pw[t] = c[0] * pw[t] + c[1] * (b0[v] - bv1);
            float sigma[3], epsr[3];
            edge_sigma_eps(g, edge_mat, q[0], q[1], q[2], sigma, epsr);
            float cb[3];
            for (int k = 0; k < 3; ++k) {
                const float dtEps = dt / epsr[k];
                cb[k] = dtEps * or_phi(sigma[k] * dtEps);
            }
            float *e = &E[4 * IDX(g, q[0], q[1], q[2])];
            e[v] = e[v] - cb[v] * pv[t];
            e[w] = e[w] + cb[w] * pw[t];
        }
        off += 4 * (size_t)cells;
    }
}

/* or_step_storage in FP32 with the convolutional PML's corrections after each update. */
void or_step_cpml(const or_grid *g, float *E, float *B, const uint32_t *edge_mat, const uint32_t *wire_tex,
                  int thin_wires, const or_source_desc *src, int n_src, int count, uint64_t *step_count,
                  const float *coef, float *psi) {
    if (count <= 0) return;
    float packed[256 * 16];
    int kept_index[256];
    int n_cells = 0, n_list = 0;
    int kept = or_set_sources(g, src, n_src, NULL, NULL, NULL, &n_cells, NULL, &n_list);
    int *cells = (int *)malloc(sizeof(int) * 4 * (size_t)(n_cells + 1));
    int *lists = (int *)malloc(sizeof(int) * (size_t)(n_list > 0 ? n_list : 1));
    kept = or_set_sources(g, src, n_src, packed, kept_index, cells, &n_cells, lists, &n_list);
    float signals[256];
    for (int k = 0; k < count; ++k) {
        const double t = ((double)(*step_count) + 0.5) * (double)g->dt;
        for (int s = 0; s < kept; ++s) signals[s] = (float)or_signal(&src[kept_index[s]], t);
        or_update_B(g, B, E, wire_tex, thin_wires);
        or_cpml_b(g, B, E, coef, psi);
        or_update_E(g, E, B, edge_mat);
        or_cpml_e(g, E, B, edge_mat, coef, psi);
        if (n_cells > 0) or_apply_sources(g, E, edge_mat, packed, signals, 0, cells, n_cells, lists);
        ++(*step_count);
    }
    free(cells);
    free(lists);
}

int or_cpml_psi_floats(const or_grid *g) {
    int total = 0, dims[3];
    for (int u = 0; u < 3; ++u) total += 4 * cpml_dims(g, u, dims);
    return total;
}

/* ------------------------------------------------------------------------------------------
 * Monitors
 * ------------------------------------------------------------------------------------------ */
static inline const float *fetch_clamped(const or_grid *g, const float *F, int x, int y, int z) {
    x = clampi(x, 0, g->nx - 1);
    y = clampi(y, 0, g->ny - 1);
    z = clampi(z, 0, g->nz - 1);
    return &F[4 * IDX(g, x, y, z)];
}

/* port_accumulate.comp edgeCurrent (shaders/port_accumulate.comp:45-50) */
static float edge_current(const or_grid *g, const float *B, int p[3], int a) {
    const int u = (a + 1) % 3, v = (a + 2) % 3;
    int pu[3] = {p[0], p[1], p[2]}, pv[3] = {p[0], p[1], p[2]};
    pu[u] -= 1;
    pv[v] -= 1;
    const float *b0 = fetch_clamped(g, B, p[0], p[1], p[2]);
    const float *bu = fetch_clamped(g, B, pu[0], pu[1], pu[2]);
    const float *bv = fetch_clamped(g, B, pv[0], pv[1], pv[2]);
    return (b0[v] - bu[v]) - (b0[u] - bv[u]);
}

void or_port_accumulate(const or_grid *g, const float *E, const float *B, const or_port *ports, int n_ports,
                        const float rot[4], float blend, int record_index, float *phasors, float *series) {
    for (int i = 0; i < n_ports; ++i) {
        const or_port *pt = &ports[i];
        int p[3] = {pt->texel[0], pt->texel[1], pt->texel[2]};
        const int a = pt->axis;
        const float s = (float)pt->sign;
        float V = 0.0f, I = 0.0f;
        if (pt->edges == 4) {
            const int u = (a + 1) % 3, v = (a + 2) % 3;
            int offs[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            offs[1][u] = 1;
            offs[2][v] = 1;
            offs[3][u] = 1;
            offs[3][v] = 1;
            for (int k = 0; k < 4; ++k) {
                int q[3] = {p[0] + offs[k][0], p[1] + offs[k][1], p[2] + offs[k][2]};
                V += 0.25f * fetch_clamped(g, E, q[0], q[1], q[2])[a];
                I += edge_current(g, B, q, a);
            }
        } else {
            V = fetch_clamped(g, E, p[0], p[1], p[2])[a];
            I = edge_current(g, B, p, a);
        }
        V *= s;
        I *= s;
        if (record_index >= 0 && series) {
            series[2 * ((size_t)record_index * (size_t)n_ports + (size_t)i) + 0] = V;
            series[2 * ((size_t)record_index * (size_t)n_ports + (size_t)i) + 1] = I;
        }
        const float obs[4] = {2.0f * rot[0] * V, -2.0f * rot[1] * V, 2.0f * rot[2] * I, -2.0f * rot[3] * I};
        float *s1 = &phasors[8 * i], *s2 = &phasors[8 * i + 4];
        for (int c = 0; c < 4; ++c) s1[c] += blend * (obs[c] - s1[c]);
        for (int c = 0; c < 4; ++c) s2[c] += blend * (s1[c] - s2[c]);
    }
}

/* PortMonitor::accumulate / FarFieldMonitor::accumulate constants
 * (src/Ports.cpp:253-262, src/FarField.cpp:331-344) */
void or_monitor_constants(double wavelength, float averaging, float dt, uint64_t step_count, float rot[4],
                          float *blend) {
    const float wl = (float)wavelength;
    const double w = 6.283185307179586476925286766559 / (double)wl;
    const double tE = (double)step_count * (double)dt;
    const double tH = tE - 0.5 * (double)dt;
    rot[0] = (float)cos(w * tE);
    rot[1] = (float)sin(w * tE);
    rot[2] = (float)cos(w * tH);
    rot[3] = (float)sin(w * tH);
    const double tau = (double)(averaging * wl);
    *blend = (float)(1.0 - exp(-(double)dt / (0.5 * tau)));
}

/* line_accumulate.comp (shaders/line_accumulate.comp:30-48): per quantity, the weighted sum of
   its terms' texel components (indices clamped to the grid; components 0-2 from E, 3-5 from B),
   kept as the instantaneous value and folded into the two-stage running DFT with the E rotation
   (kind 0) or the B rotation (kind 1). quantities: 4 ints each (first, count, kind, unused);
   terms: 4 ints each (x, y, z, component). */
void or_line_accumulate(const or_grid *g, const float *E, const float *B, const int *quantities, int n,
                        const int *terms, const float *weights, const float rot[4], float blend, float *phasors,
                        float *now) {
    for (int q = 0; q < n; ++q) {
        const int first = quantities[4 * q], count = quantities[4 * q + 1], kind = quantities[4 * q + 2];
        float sum = 0.0f;
        for (int k = first; k < first + count; ++k) {
            const int *t = terms + 4 * k;
            const int x = t[0] < 0 ? 0 : (t[0] > g->nx - 1 ? g->nx - 1 : t[0]);
            const int y = t[1] < 0 ? 0 : (t[1] > g->ny - 1 ? g->ny - 1 : t[1]);
            const int z = t[2] < 0 ? 0 : (t[2] > g->nz - 1 ? g->nz - 1 : t[2]);
            const float *F = t[3] < 3 ? E : B;
            const size_t i = ((size_t)z * g->ny + y) * g->nx + x;
            sum += weights[k] * F[4 * i + t[3] % 3];
        }
        now[q] = sum;
        const float rc = kind == 0 ? rot[0] : rot[2], rs = kind == 0 ? rot[1] : rot[3];
        const float two = 2.0f * sum;
        const float ox = two * rc, oy = two * -rs;
        float *ph = phasors + 4 * q;
        ph[0] += blend * (ox - ph[0]);
        ph[1] += blend * (oy - ph[1]);
        ph[2] += blend * (ph[0] - ph[2]);
        ph[3] += blend * (ph[1] - ph[3]);
    }
}

/* ntff_common.glsl ntffPoint (shaders/ntff_common.glsl:30-53) */
void or_ntff_point(const int bmin[3], const int bmax[3], const int fo[7], int index, int node[3], float normal[3],
                   float *weight) {
    int f = 0;
    for (int i = 1; i < 6; ++i)
        if (index >= fo[i]) f = i;
    const int a = f >> 1, side = f & 1;
    const int u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
    const int nu = bmax[u] - bmin[u], nv = bmax[v] - bmin[v];
    const int local = index - fo[f];
    const int iu = local % (nu + 1), iv = local / (nu + 1);
    node[0] = bmin[0];
    node[1] = bmin[1];
    node[2] = bmin[2];
    node[a] = side == 1 ? bmax[a] : bmin[a];
    node[u] = bmin[u] + iu;
    node[v] = bmin[v] + iv;
    *weight = ((iu == 0 || iu == nu) ? 0.5f : 1.0f) * ((iv == 0 || iv == nv) ? 0.5f : 1.0f);
    normal[0] = normal[1] = normal[2] = 0.0f;
    normal[a] = side == 1 ? 1.0f : -1.0f;
}

int or_ntff_face_offsets(const int bmin[3], const int bmax[3], int fo[7]) {
    fo[0] = 0;
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        fo[f + 1] = fo[f] + (bmax[u] - bmin[u] + 1) * (bmax[v] - bmin[v] + 1);
    }
    return fo[6];
}

/* field_sampling.glsl eAtNode / bAtNode (shaders/field_sampling.glsl:17-32) */
void or_e_at_node(const or_grid *g, const float *E, int x, int y, int z, float out[3]) {
    const float *e0 = fetch_clamped(g, E, x, y, z);
    out[0] = 0.5f * (e0[0] + fetch_clamped(g, E, x - 1, y, z)[0]);
    out[1] = 0.5f * (e0[1] + fetch_clamped(g, E, x, y - 1, z)[1]);
    out[2] = 0.5f * (e0[2] + fetch_clamped(g, E, x, y, z - 1)[2]);
}

void or_b_at_node(const or_grid *g, const float *B, int x, int y, int z, float out[3]) {
    const float *b0 = fetch_clamped(g, B, x, y, z);
    const float bx = b0[0] + fetch_clamped(g, B, x, y - 1, z)[0] + fetch_clamped(g, B, x, y, z - 1)[0] +
                     fetch_clamped(g, B, x, y - 1, z - 1)[0];
    const float by = b0[1] + fetch_clamped(g, B, x - 1, y, z)[1] + fetch_clamped(g, B, x, y, z - 1)[1] +
                     fetch_clamped(g, B, x - 1, y, z - 1)[1];
    const float bz = b0[2] + fetch_clamped(g, B, x - 1, y, z)[2] + fetch_clamped(g, B, x, y - 1, z)[2] +
                     fetch_clamped(g, B, x - 1, y - 1, z)[2];
    out[0] = 0.25f * bx;
    out[1] = 0.25f * by;
    out[2] = 0.25f * bz;
}

/* ntff_accumulate.comp (shaders/ntff_accumulate.comp:30-52) */
void or_ntff_accumulate(const or_grid *g, const float *E, const float *B, const int bmin[3], const int bmax[3],
                        const int fo[7], int n_points, const float rot[4], float blend, float *ph) {
    for (int i = 0; i < n_points; ++i) {
        int node[3];
        float n[3], w;
        or_ntff_point(bmin, bmax, fo, i, node, n, &w);
        float Ev[3], Hv[3];
        or_e_at_node(g, E, node[0], node[1], node[2], Ev);
        or_b_at_node(g, B, node[0], node[1], node[2], Hv);
        float *p = &ph[32 * (size_t)i];
        float *s1eRe = p + 0, *s1eIm = p + 4, *s1hRe = p + 8, *s1hIm = p + 12;
        float *s2eRe = p + 16, *s2eIm = p + 20, *s2hRe = p + 24, *s2hIm = p + 28;
        const float ce = 2.0f * rot[0], se = -2.0f * rot[1], ch = 2.0f * rot[2], sh = -2.0f * rot[3];
        for (int c = 0; c < 3; ++c) {
            s1eRe[c] += blend * (ce * Ev[c] - s1eRe[c]);
            s1eIm[c] += blend * (se * Ev[c] - s1eIm[c]);
            s1hRe[c] += blend * (ch * Hv[c] - s1hRe[c]);
            s1hIm[c] += blend * (sh * Hv[c] - s1hIm[c]);
        }
        for (int c = 0; c < 3; ++c) {
            s2eRe[c] += blend * (s1eRe[c] - s2eRe[c]);
            s2eIm[c] += blend * (s1eIm[c] - s2eIm[c]);
            s2hRe[c] += blend * (s1hRe[c] - s2hRe[c]);
            s2hIm[c] += blend * (s1hIm[c] - s2hIm[c]);
        }
    }
}

/* The Huygens-node samples that ntff_accumulate.comp forms (shaders/ntff_accumulate.comp:30-40),
 * without the exponential means: out[6 i + c] = E at point i, out[6 i + 3 + c] = H. The running
 * DFTs of these samples (a library extension with no reference counterpart) are summed by the
 * caller in FP64. */
void or_ntff_samples(const or_grid *g, const float *E, const float *B, const int bmin[3], const int bmax[3],
                     const int fo[7], int n_points, float *out) {
    for (int i = 0; i < n_points; ++i) {
        int node[3];
        float n[3], w;
        or_ntff_point(bmin, bmax, fo, i, node, n, &w);
        or_e_at_node(g, E, node[0], node[1], node[2], out + 6 * (size_t)i);
        or_b_at_node(g, B, node[0], node[1], node[2], out + 6 * (size_t)i + 3);
    }
}

static inline void crossf(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - b[1] * a[2];
    o[1] = a[2] * b[0] - b[2] * a[0];
    o[2] = a[0] * b[1] - b[0] * a[1];
}
static inline float dotf(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ntff_pattern.comp + ntff_reduce.comp + FarFieldMonitor::evaluate chunking
 * (shaders/ntff_pattern.comp:45-110, shaders/ntff_reduce.comp:19-39, src/FarField.cpp:226-243,256-276,294-297) */
void or_ntff_eval_f32(const int bmin[3], const int bmax[3], const int fo[7], int n_points, const float *ph,
                      const float *dirs, int n_dirs, float k, const float center[3], float *out) {
    int slices = (n_points + 4095) / 4096;
    slices = slices < 1 ? 1 : (slices > 64 ? 64 : slices);
    int chunk = (int)(1.5e8 / (double)(n_points > 1 ? n_points : 1)) / 128 * 128;
    chunk = chunk < 128 ? 128 : (chunk > 4096 ? 4096 : chunk);
    const int per = (n_points + slices - 1) / slices;
    /* per-point currents and positions as the tile loader builds them */
    float *pos = (float *)malloc(sizeof(float) * 3 * (size_t)n_points);
    float *cur = (float *)malloc(sizeof(float) * 12 * (size_t)n_points);
    float *pw = (float *)malloc(sizeof(float) * (size_t)n_points);
    for (int i = 0; i < n_points; ++i) {
        int node[3];
        float n[3], w;
        or_ntff_point(bmin, bmax, fo, i, node, n, &w);
        const float *s2 = &ph[32 * (size_t)i + 16];
        for (int a = 0; a < 3; ++a) pos[3 * i + a] = (float)node[a] - center[a];
        float t[3];
        crossf(n, s2 + 8, t);  for (int a = 0; a < 3; ++a) cur[12 * i + 0 + a] = w * t[a];
        crossf(n, s2 + 12, t); for (int a = 0; a < 3; ++a) cur[12 * i + 3 + a] = w * t[a];
        crossf(n, s2 + 0, t);  for (int a = 0; a < 3; ++a) cur[12 * i + 6 + a] = -w * t[a];
        crossf(n, s2 + 4, t);  for (int a = 0; a < 3; ++a) cur[12 * i + 9 + a] = -w * t[a];
        float c1[3], c2[3];
        crossf(s2 + 0, s2 + 8, c1);
        crossf(s2 + 4, s2 + 12, c2);
        const float sum[3] = {c1[0] + c2[0], c1[1] + c2[1], c1[2] + c2[2]};
        pw[i] = w * dotf(sum, n);
    }
    float *part = (float *)malloc(sizeof(float) * 12 * (size_t)slices);
    float *pslice = (float *)malloc(sizeof(float) * (size_t)slices);
    for (int first = 0; first < n_dirs; first += chunk) {
        const int count = n_dirs - first < chunk ? n_dirs - first : chunk;
        const int power_chunk = first == 0;
        if (power_chunk) {
            /* slice power: each of the 128 threads sums points lid, lid+128, ... then a tree */
            float sP[128];
            for (int sl = 0; sl < slices; ++sl) {
                const int f0 = sl * per, last = f0 + per < n_points ? f0 + per : n_points;
                for (int lid = 0; lid < 128; ++lid) {
                    float P = 0.0f;
                    for (int base = f0; base < last; base += 128) {
                        const int i = base + lid;
                        if (i < last) P += pw[i];
                    }
                    sP[lid] = P;
                }
                for (int stride = 64; stride > 0; stride >>= 1)
                    for (int lid = 0; lid < stride; ++lid) sP[lid] += sP[lid + stride];
                pslice[sl] = sP[0];
            }
        }
        for (int dl = 0; dl < count; ++dl) {
            const int d = first + dl;
            const float r[3] = {dirs[4 * d + 0], dirs[4 * d + 1], dirs[4 * d + 2]};
            for (int sl = 0; sl < slices; ++sl) {
                const int f0 = sl * per, last = f0 + per < n_points ? f0 + per : n_points;
                float Nr[3] = {0, 0, 0}, Ni[3] = {0, 0, 0}, Lr[3] = {0, 0, 0}, Li[3] = {0, 0, 0};
                for (int i = f0; i < last; ++i) {
                    const float psi = k * dotf(r, &pos[3 * i]);
                    const float c = cosf(psi), s = sinf(psi);
                    const float *q = &cur[12 * i];
                    for (int a = 0; a < 3; ++a) {
                        Nr[a] += q[0 + a] * c - q[3 + a] * s;
                        Ni[a] += q[0 + a] * s + q[3 + a] * c;
                        Lr[a] += q[6 + a] * c - q[9 + a] * s;
                        Li[a] += q[6 + a] * s + q[9 + a] * c;
                    }
                }
                for (int a = 0; a < 3; ++a) {
                    part[12 * sl + a] = Nr[a];
                    part[12 * sl + 3 + a] = Ni[a];
                    part[12 * sl + 6 + a] = Lr[a];
                    part[12 * sl + 9 + a] = Li[a];
                }
            }
            float Nr[3] = {0, 0, 0}, Ni[3] = {0, 0, 0}, Lr[3] = {0, 0, 0}, Li[3] = {0, 0, 0}, P = 0.0f;
            for (int sl = 0; sl < slices; ++sl) {
                for (int a = 0; a < 3; ++a) {
                    Nr[a] += part[12 * sl + a];
                    Ni[a] += part[12 * sl + 3 + a];
                    Lr[a] += part[12 * sl + 6 + a];
                    Li[a] += part[12 * sl + 9 + a];
                }
                P += (power_chunk && dl == 0) ? pslice[sl] : 0.0f;
            }
            float Fr[3], Fi[3], cr[3], ci[3];
            const float dr = dotf(Nr, r), di = dotf(Ni, r);
            crossf(r, Lr, cr);
            crossf(r, Li, ci);
            for (int a = 0; a < 3; ++a) {
                Fr[a] = (Nr[a] - dr * r[a]) - cr[a];
                Fi[a] = (Ni[a] - di * r[a]) - ci[a];
            }
            out[2 * d] = dotf(Fr, Fr) + dotf(Fi, Fi);
            out[2 * d + 1] = 0.5f * P;
        }
    }
    free(pos);
    free(cur);
    free(pw);
    free(part);
    free(pslice);
}

void or_ntff_eval_f64(const int bmin[3], const int bmax[3], const int fo[7], int n_points, const float *ph,
                      const double *dirs, int n_dirs, double k, const double center[3], double *outF2, double *Pout) {
    double P = 0.0;
    for (int i = 0; i < n_points; ++i) {
        int node[3];
        float nf[3], wf;
        or_ntff_point(bmin, bmax, fo, i, node, nf, &wf);
        const float *s2 = &ph[32 * (size_t)i + 16];
        const double n[3] = {nf[0], nf[1], nf[2]}, w = wf;
        const double eR[3] = {s2[0], s2[1], s2[2]}, eI[3] = {s2[4], s2[5], s2[6]};
        const double hR[3] = {s2[8], s2[9], s2[10]}, hI[3] = {s2[12], s2[13], s2[14]};
        const double c1[3] = {eR[1] * hR[2] - eR[2] * hR[1], eR[2] * hR[0] - eR[0] * hR[2], eR[0] * hR[1] - eR[1] * hR[0]};
        const double c2[3] = {eI[1] * hI[2] - eI[2] * hI[1], eI[2] * hI[0] - eI[0] * hI[2], eI[0] * hI[1] - eI[1] * hI[0]};
        P += w * ((c1[0] + c2[0]) * n[0] + (c1[1] + c2[1]) * n[1] + (c1[2] + c2[2]) * n[2]);
    }
    *Pout = 0.5 * P;
    for (int d = 0; d < n_dirs; ++d) {
        const double *r = &dirs[3 * d];
        double Nr[3] = {0, 0, 0}, Ni[3] = {0, 0, 0}, Lr[3] = {0, 0, 0}, Li[3] = {0, 0, 0};
        for (int i = 0; i < n_points; ++i) {
            int node[3];
            float nf[3], wf;
            or_ntff_point(bmin, bmax, fo, i, node, nf, &wf);
            const float *s2 = &ph[32 * (size_t)i + 16];
            const double n[3] = {nf[0], nf[1], nf[2]}, w = wf;
            const double q[3] = {node[0] - center[0], node[1] - center[1], node[2] - center[2]};
            const double psi = k * (r[0] * q[0] + r[1] * q[1] + r[2] * q[2]);
            const double c = cos(psi), s = sin(psi);
            const double hR[3] = {s2[8], s2[9], s2[10]}, hI[3] = {s2[12], s2[13], s2[14]};
            const double eR[3] = {s2[0], s2[1], s2[2]}, eI[3] = {s2[4], s2[5], s2[6]};
            double jr[3], ji[3], mr[3], mi[3];
            jr[0] = w * (n[1] * hR[2] - n[2] * hR[1]); jr[1] = w * (n[2] * hR[0] - n[0] * hR[2]); jr[2] = w * (n[0] * hR[1] - n[1] * hR[0]);
            ji[0] = w * (n[1] * hI[2] - n[2] * hI[1]); ji[1] = w * (n[2] * hI[0] - n[0] * hI[2]); ji[2] = w * (n[0] * hI[1] - n[1] * hI[0]);
            mr[0] = -w * (n[1] * eR[2] - n[2] * eR[1]); mr[1] = -w * (n[2] * eR[0] - n[0] * eR[2]); mr[2] = -w * (n[0] * eR[1] - n[1] * eR[0]);
            mi[0] = -w * (n[1] * eI[2] - n[2] * eI[1]); mi[1] = -w * (n[2] * eI[0] - n[0] * eI[2]); mi[2] = -w * (n[0] * eI[1] - n[1] * eI[0]);
            for (int a = 0; a < 3; ++a) {
                Nr[a] += jr[a] * c - ji[a] * s;
                Ni[a] += jr[a] * s + ji[a] * c;
                Lr[a] += mr[a] * c - mi[a] * s;
                Li[a] += mr[a] * s + mi[a] * c;
            }
        }
        const double dr = Nr[0] * r[0] + Nr[1] * r[1] + Nr[2] * r[2];
        const double di = Ni[0] * r[0] + Ni[1] * r[1] + Ni[2] * r[2];
        const double cr[3] = {r[1] * Lr[2] - r[2] * Lr[1], r[2] * Lr[0] - r[0] * Lr[2], r[0] * Lr[1] - r[1] * Lr[0]};
        const double ci[3] = {r[1] * Li[2] - r[2] * Li[1], r[2] * Li[0] - r[0] * Li[2], r[0] * Li[1] - r[1] * Li[0]};
        double F2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            const double fr = (Nr[a] - dr * r[a]) - cr[a], fi = (Ni[a] - di * r[a]) - ci[a];
            F2 += fr * fr + fi * fi;
        }
        outF2[d] = F2;
    }
}

/* FarFieldMonitor::buildDirections + boresightFrame (src/FarField.cpp:201-257, src/FarField.hpp:29-37) */
int or_build_directions(const float axis[3], float *dirs, int cap, int *map_index, float frame[9]) {
    int nd = 0;
    const float cuts[3][2][3] = {{{1, 0, 0}, {0, 1, 0}}, {{1, 0, 0}, {0, 0, 1}}, {{0, 1, 0}, {0, 0, 1}}};
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 180; ++i) {
            const float deg = 360.0f * (float)i / 180.0f;
            const float rad = deg * (3.14159265358979323846f / 180.0f);
            const float cs = cosf(rad), sn = sinf(rad);
            if (nd < cap) {
                for (int a = 0; a < 3; ++a) dirs[4 * nd + a] = cs * cuts[c][0][a] + sn * cuts[c][1][a];
                dirs[4 * nd + 3] = 0.0f;
            }
            ++nd;
        }
    int ax = 2;
    for (int i = 0; i < 3; ++i)
        if (fabsf(axis[i]) > fabsf(axis[ax])) ax = i;
    float n[3] = {0, 0, 0}, eu[3], ev[3];
    n[ax] = axis[ax] >= 0.0f ? 1.0f : -1.0f;
    if (fabsf(n[2]) > 0.5f) {
        eu[0] = 1; eu[1] = 0; eu[2] = 0;
        ev[0] = 0; ev[1] = 1; ev[2] = 0;
    } else {
        ev[0] = 0; ev[1] = 0; ev[2] = 1;
        crossf(n, ev, eu);
    }
    for (int a = 0; a < 3; ++a) {
        frame[a] = n[a];
        frame[3 + a] = eu[a];
        frame[6 + a] = ev[a];
    }
    for (int i = 0; i < 2 * 81 * 81; ++i) map_index[i] = -1;
    for (int h = 0; h < 2; ++h)
        for (int r = 0; r < 81; ++r)
            for (int c = 0; c < 81; ++c) {
                const float u = -1.0f + 2.0f * (float)c / 80.0f;
                const float v = 1.0f - 2.0f * (float)r / 80.0f;
                const float rho2 = u * u + v * v;
                if (rho2 > 1.0f) continue;
                const float w = sqrtf(fmaxf_(1.0f - rho2, 0.0f));
                map_index[(h * 81 + r) * 81 + c] = nd;
                const float ww = h == 0 ? w : -w;
                if (nd < cap) {
                    for (int a = 0; a < 3; ++a) dirs[4 * nd + a] = u * eu[a] + v * ev[a] + ww * n[a];
                    dirs[4 * nd + 3] = 0.0f;
                }
                ++nd;
            }
    return nd;
}

/* poynting_flux.comp + FieldAnalyzer::planeFlux (shaders/poynting_flux.comp:32-63,
 * src/FieldAnalyzer.cpp:317-359) */
double or_plane_flux(const or_grid *g, const float *E, const float *B, int axis, int index, int rmin_u, int rmin_v,
                     int rmax_u, int rmax_v) {
    const int size[3] = {g->nx, g->ny, g->nz};
    if (index < 0 || index >= size[axis]) return 0.0;
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
    if (rmin_u < 0) rmin_u = 0;
    if (rmin_v < 0) rmin_v = 0;
    if (rmax_u > size[u]) rmax_u = size[u];
    if (rmax_v > size[v]) rmax_v = size[v];
    const int su = rmax_u - rmin_u, sv = rmax_v - rmin_v;
    if (su <= 0 || sv <= 0) return 0.0;
    const int gx = (su + 15) / 16, gy = (sv + 15) / 16;
    double sum = 0.0;
    for (int by = 0; by < gy; ++by)
        for (int bx = 0; bx < gx; ++bx) {
            float s[256];
            for (int ly = 0; ly < 16; ++ly)
                for (int lx = 0; lx < 16; ++lx) {
                    const int lu = bx * 16 + lx, lv = by * 16 + ly;
                    float flux = 0.0f;
                    if (lu < su && lv < sv) {
                        const int uu = rmin_u + lu, vv = rmin_v + lv;
                        int p[3];
                        if (axis == 0) { p[0] = index; p[1] = uu; p[2] = vv; }
                        else if (axis == 1) { p[0] = uu; p[1] = index; p[2] = vv; }
                        else { p[0] = uu; p[1] = vv; p[2] = index; }
                        float e[3], b[3], S[3];
                        or_e_at_node(g, E, p[0], p[1], p[2], e);
                        or_b_at_node(g, B, p[0], p[1], p[2], b);
                        crossf(e, b, S);
                        flux = S[axis];
                    }
                    s[ly * 16 + lx] = flux;
                }
            for (int stride = 128; stride > 0; stride >>= 1)
                for (int l = 0; l < stride; ++l) s[l] += s[l + stride];
            sum += (double)s[0];
        }
    return sum;
}

double or_box_flux(const or_grid *g, const float *E, const float *B, const int bmin[3], const int bmax[3]) {
    int mn[3], mx[3];
    for (int a = 0; a < 3; ++a) {
        mn[a] = bmin[a] < bmax[a] ? bmin[a] : bmax[a];
        mx[a] = bmin[a] < bmax[a] ? bmax[a] : bmin[a];
    }
    double total = 0.0;
    for (int axis = 0; axis < 3; ++axis) {
        const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
        total -= or_plane_flux(g, E, B, axis, mn[axis], mn[u], mn[v], mx[u] + 1, mx[v] + 1);
        total += or_plane_flux(g, E, B, axis, mx[axis], mn[u], mn[v], mx[u] + 1, mx[v] + 1);
    }
    return total;
}

/* probe_gather.comp + sampleBlock + collect (shaders/probe_gather.comp:20-26, src/FieldAnalyzer.cpp:19-40,
 * 274-303, 367-388) */
static float sample_block(const float *blk, const int origin[3], const int extent[3], const float q[3], int ch) {
    const float fl[3] = {floorf(q[0]), floorf(q[1]), floorf(q[2])};
    const float f[3] = {q[0] - fl[0], q[1] - fl[1], q[2] - fl[2]};
    const int i0[3] = {(int)fl[0] - origin[0], (int)fl[1] - origin[1], (int)fl[2] - origin[2]};
    float result = 0.0f;
    for (int dz = 0; dz <= 1; ++dz)
        for (int dy = 0; dy <= 1; ++dy)
            for (int dx = 0; dx <= 1; ++dx) {
                const float w = (dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
                if (w != 0.0f) {
                    const int x = clampi(i0[0] + dx, 0, extent[0] - 1), y = clampi(i0[1] + dy, 0, extent[1] - 1),
                              z = clampi(i0[2] + dz, 0, extent[2] - 1);
                    result += w * blk[4 * ((z * 3 + y) * 3 + x) + ch];
                }
            }
    return result;
}

void or_probe(const or_grid *g, const float *E, const float *B, const float lp[3], float out[11]) {
    const int size[3] = {g->nx, g->ny, g->nz};
    float pos[3];
    int origin[3], extent[3];
    for (int a = 0; a < 3; ++a) {
        const float hi = (float)(size[a] - 1);
        pos[a] = lp[a] < 0.0f ? 0.0f : (lp[a] > hi ? hi : lp[a]);
        const int mx = size[a] - 3 > 0 ? size[a] - 3 : 0;
        origin[a] = clampi((int)floorf(pos[a]) - 1, 0, mx);
        extent[a] = size[a] - origin[a] < 3 ? size[a] - origin[a] : 3;
    }
    float eb[27 * 4], bb[27 * 4];
    for (int t = 0; t < 27; ++t) {
        const int x = clampi(origin[0] + t % 3, 0, g->nx - 1), y = clampi(origin[1] + (t / 3) % 3, 0, g->ny - 1),
                  z = clampi(origin[2] + t / 9, 0, g->nz - 1);
        memcpy(&eb[4 * t], &E[4 * IDX(g, x, y, z)], 16);
        memcpy(&bb[4 * t], &B[4 * IDX(g, x, y, z)], 16);
    }
    const float qE[3][3] = {{pos[0] - 0.5f, pos[1] - 0.0f, pos[2] - 0.0f},
                            {pos[0] - 0.0f, pos[1] - 0.5f, pos[2] - 0.0f},
                            {pos[0] - 0.0f, pos[1] - 0.0f, pos[2] - 0.5f}};
    const float qB[3][3] = {{pos[0] - 0.0f, pos[1] - 0.5f, pos[2] - 0.5f},
                            {pos[0] - 0.5f, pos[1] - 0.0f, pos[2] - 0.5f},
                            {pos[0] - 0.5f, pos[1] - 0.5f, pos[2] - 0.0f}};
    float e[3], b[3], S[3];
    for (int c = 0; c < 3; ++c) {
        e[c] = sample_block(eb, origin, extent, qE[c], c);
        b[c] = sample_block(bb, origin, extent, qB[c], c);
    }
    crossf(e, b, S);
    int cell[3];
    for (int a = 0; a < 3; ++a) cell[a] = clampi((int)floorf(pos[a]) - origin[a], 0, extent[a] - 1);
    const int ci = (cell[2] * 3 + cell[1]) * 3 + cell[0];
    for (int a = 0; a < 3; ++a) {
        out[a] = e[a];
        out[3 + a] = b[a];
        out[6 + a] = S[a];
    }
    out[9] = eb[4 * ci + 3];
    out[10] = bb[4 * ci + 3];
}

/* PortMonitor::evaluateSweep (src/Ports.cpp:187-278) */
double or_sweep(const float *series, size_t n, int n_ports, int port, double dt, double wl_min, double wl_max,
                int points, double z0, double *out, double *band) {
    const double kEta0 = 376.730313668;
    const double TWO_PI = 6.283185307179586476925286766559;
    if (points < 2) points = 2;
    const double lmax = wl_max > wl_min ? wl_max : wl_min, lmin = wl_max < wl_min ? wl_max : wl_min;
    const double nuLo = 1.0 / lmax, nuHi = 1.0 / lmin;
    double peak = 0.0, late = 0.0;
    size_t tail = n / 20 > 1 ? n / 20 : 1;
    for (size_t k = 0; k < n; ++k) {
        const double v = fabs((double)series[2 * (k * (size_t)n_ports)]);
        if (v > peak) peak = v;
        if (k >= n - tail && v > late) late = v;
    }
    const double residual = peak > 0.0 ? late / peak : 1.0;
    if (z0 < 1e-6) z0 = 1e-6;
    for (int f = 0; f < points; ++f) {
        const double nu = nuLo + (nuHi - nuLo) * f / (points - 1);
        const double w = TWO_PI * nu;
        const double sr = cos(-w * dt), si = sin(-w * dt);
        double rr = 1.0, ri = 0.0, vr = 0.0, vi = 0.0, ir = 0.0, ii = 0.0;
        for (size_t k = 0; k < n; ++k) {
            const double V = (double)series[2 * (k * (size_t)n_ports + (size_t)port)];
            const double I = (double)series[2 * (k * (size_t)n_ports + (size_t)port) + 1];
            vr += V * rr; vi += V * ri;
            ir += I * rr; ii += I * ri;
            const double nr = rr * sr - ri * si, ni = rr * si + ri * sr;
            rr = nr; ri = ni;
            if ((k & 1023) == 1023) {
                const double m = hypot(rr, ri);
                rr /= m; ri /= m;
            }
        }
        { /* sumI *= polar(1, w dt / 2) */
            const double hr = cos(w * 0.5 * dt), hi = sin(w * 0.5 * dt);
            const double nr = ir * hr - ii * hi, ni = ir * hi + ii * hr;
            ir = nr; ii = ni;
        }
        double *o = &out[5 * f];
        o[0] = 1.0 / nu;
        o[1] = o[2] = o[3] = o[4] = 0.0;
        if (hypot(ir, ii) > 1e-30) {
            /* z = -sumV / sumI * eta0 */
            const double den = ir * ir + ii * ii;
            const double qr = (vr * ir + vi * ii) / den, qi = (vi * ir - vr * ii) / den;
            const double zr = -qr * kEta0, zi = -qi * kEta0;
            /* gamma = (z - z0) / (z + z0) */
            const double ar = zr - z0, br = zr + z0;
            const double gd = br * br + zi * zi;
            const double gr = (ar * br + zi * zi) / gd, gi = (zi * br - ar * zi) / gd;
            double gm = hypot(gr, gi);
            o[1] = zr;
            o[2] = zi;
            o[3] = gm > 1e-12 ? 20.0 * log10(gm) : -200.0;
            if (gm > 0.999999) gm = 0.999999;
            o[4] = (1.0 + gm) / (1.0 - gm);
        }
    }
    size_t best = 0;
    for (int i = 1; i < points; ++i)
        if (out[5 * i + 3] < out[5 * best + 3]) best = (size_t)i;
    double nuBest = 1.0 / out[5 * best + 0];
    if (best > 0 && best + 1 < (size_t)points) {
        const double a = out[5 * (best - 1) + 3], b = out[5 * best + 3], c = out[5 * (best + 1) + 3];
        const double denom = a - 2.0 * b + c;
        const double dnu = (nuHi - nuLo) / (points - 1);
        if (fabs(denom) > 1e-12) {
            double t = 0.5 * (a - c) / denom;
            t = t < -1.0 ? -1.0 : (t > 1.0 ? 1.0 : t);
            nuBest += t * dnu;
        }
    }
    band[0] = (double)best;
    band[1] = 1.0 / nuBest;
    band[2] = 0.0;
    band[3] = band[4] = 0.0;
    if (out[5 * best + 3] < -10.0) {
        double edges[2];
        for (int e = 0; e < 2; ++e) {
            const int dir = e == 0 ? -1 : 1;
            int i = (int)best;
            while (i + dir >= 0 && i + dir < points && out[5 * (i + dir) + 3] < -10.0) i += dir;
            const int j = i + dir;
            const double nuI = 1.0 / out[5 * i + 0];
            if (j < 0 || j >= points) {
                edges[e] = nuI;
                continue;
            }
            const double nuJ = 1.0 / out[5 * j + 0];
            const double sI = out[5 * i + 3], sJ = out[5 * j + 3];
            edges[e] = nuI + (nuJ - nuI) * (-10.0 - sI) / (sJ - sI);
        }
        band[2] = 1.0;
        band[3] = 1.0 / edges[0];
        band[4] = 1.0 / edges[1];
    }
    return residual;
}
