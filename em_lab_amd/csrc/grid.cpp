// libemfdtd.so host side: context, SimulationGrid entry points, materials and sources.
//
// Host logic restates EM_Lab's SimulationGrid (src/SimulationGrid.cpp) where the result must
// match bit for bit (clamps, edge averaging, binary16 quantisation, source lists, FP64 signals);
// the GPU state and launch sequence are this library's own design.
#include "context.h"
#include "expf.h"
#include "host_common.h"
#include "signal.h"

#include <algorithm>
#include <cstdint>
#include <array>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace emfd {

namespace {
thread_local std::string g_error;
constexpr float kMaxCourant = 0.99f / 1.7320508f;  // src/SimulationGrid.cpp:24
constexpr float kPerfectEdgeSigma = 6.0e4f;        // src/SimulationGrid.hpp:103
constexpr int kMaxSources = 256;                   // src/SimulationGrid.hpp:157
}  // namespace

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? EMFD_ERR_NOMEM : EMFD_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

int DevBuf::ensure(size_t need) {
    if (need <= bytes && ptr) return EMFD_OK;
    release();
    if (need == 0) need = 16;
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) {
        ptr = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    bytes = need;
    return EMFD_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

uint16_t f32_to_f16(float f) {
    // binary16, round to nearest even: glm::packHalf2x16 (src/SimulationGrid.cpp:260-261)
    const _Float16 h = static_cast<_Float16>(f);
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

// SourceDesc::signal (src/SimulationGrid.cpp:99-130), double precision
double source_signal(const emfd_source_desc &s, double tAbs) { return source_signal_d(s, tAbs); }

// The convolutional PML's auxiliary fields back to 0 (no-op with the graded absorber).
int clear_psi(emfd_ctx *ctx) {
    if (ctx->absorber != EMFD_ABSORBER_CPML || !ctx->d_psi.ptr) return EMFD_OK;
    EMFD_HIP(launch_fill(ctx->d_psi.as<float>(), ctx->d_psi.bytes / sizeof(float), 0.0f, ctx->stream));
    return EMFD_OK;
}

emfd_grid_config clamp_config(const emfd_grid_config &requested) {
    emfd_grid_config c = requested;
    c.pml_thickness = std::clamp(c.pml_thickness, 0, 64);
    const int min_size = 2 * c.pml_thickness + 8;
    for (int a = 0; a < 3; ++a) c.size[a] = std::max(c.size[a], min_size);
    c.courant = std::clamp(c.courant, 0.05f, kMaxCourant);
    c.pml_sigma_max = std::max(c.pml_sigma_max, 0.0f);
    c.pml_grading = std::clamp(c.pml_grading, 0.25f, 12.0f);
    return c;
}

std::vector<float> absorber_profiles(const emfd_grid_config &c, int *stride_out) {
    // evaluated in float exactly as the shader does
    const int stride = std::max({c.size[0], c.size[1], c.size[2]});
    std::vector<float> prof(static_cast<size_t>(6) * stride, 0.0f);
    const float L = static_cast<float>(c.pml_thickness);
    auto profile = [&](float depth) -> float {
        if (depth <= 0.0f) return 0.0f;
        const float u = std::min(depth / static_cast<float>(c.pml_thickness), 1.0f);
        return c.pml_sigma_max * (det_expf(c.pml_grading * u) - 1.0f) / (det_expf(c.pml_grading) - 1.0f);
    };
    for (int a = 0; a < 3; ++a) {
        const float hi = static_cast<float>(c.size[a] - 1) - L;
        for (int h = 0; h < 2; ++h)
            for (int i = 0; i < c.size[a]; ++i) {
                const float x = static_cast<float>(i) + (h ? 0.5f : 0.0f);
                prof[static_cast<size_t>(2 * a + h) * stride + i] = profile(std::max(L - x, x - hi));
            }
    }
    *stride_out = stride;
    return prof;
}

ClassKey ClassMap::vacuum() {
    const uint16_t one = f32_to_f16(1.0f);
    return {0, 0, 0, one, one, one, one, one, one, one, one, one};
}

ClassMap::ClassMap() {
    map.emplace(vacuum(), 0u);
    keys.push_back(vacuum());
}

bool absorber_has_materials(const emfd_grid_config &c, const uint32_t *ids) {
    const int nx = c.size[0], ny = c.size[1], nz = c.size[2], L = c.pml_thickness;
    const auto edge = [L](int i, int n) { return i < L || i >= n - (L + 1); };
    size_t p = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            const bool yz = edge(y, ny) || edge(z, nz);
            for (int x = 0; x < nx; ++x, ++p)
                if (ids[p] != 0u && (yz || edge(x, nx))) return true;
        }
    return false;
}

void classify_cells(const emfd_grid_config &cfg, const float *sigma, const float *eps_r,
                    const emfd_edge_override *edges, int64_t n_edges, const emfd_wire_coefficient *wires,
                    int64_t n_wires, ClassMap &cm, uint32_t *ids) {
    const int nx = cfg.size[0], ny = cfg.size[1], nz = cfg.size[2];
    // sub-cell overrides sorted by texel, keeping their order (later entries win, as the
    // reference's sequential loop at src/SimulationGrid.cpp:251-254)
    struct EdgeOv { size_t idx; int axis; float sigma; };
    std::vector<EdgeOv> ov;
    for (int64_t k = 0; k < n_edges; ++k) {
        const emfd_edge_override &o = edges[k];
        if (o.texel[0] < 0 || o.texel[1] < 0 || o.texel[2] < 0 || o.texel[0] >= nx || o.texel[1] >= ny ||
            o.texel[2] >= nz || o.axis < 0 || o.axis > 2)
            continue;
        ov.push_back({(static_cast<size_t>(o.texel[2]) * ny + o.texel[1]) * nx + o.texel[0], o.axis,
                      std::max(o.sigma, 0.0f)});
    }
    std::stable_sort(ov.begin(), ov.end(), [](const EdgeOv &a, const EdgeOv &b) { return a.idx < b.idx; });
    struct WireOv { size_t idx; int slot; float factor; };
    std::vector<WireOv> wv;
    for (int64_t k = 0; k < n_wires; ++k) {
        const emfd_wire_coefficient &w = wires[k];
        if (w.texel[0] < 0 || w.texel[1] < 0 || w.texel[2] < 0 || w.texel[0] >= nx || w.texel[1] >= ny ||
            w.texel[2] >= nz || w.component < 0 || w.component > 2)
            continue;
        wv.push_back({(static_cast<size_t>(w.texel[2]) * ny + w.texel[1]) * nx + w.texel[0],
                      2 * w.component + (w.term == 0 ? 0 : 1), w.factor});
    }
    std::stable_sort(wv.begin(), wv.end(), [](const WireOv &a, const WireOv &b) { return a.idx < b.idx; });
    // edge averages (src/SimulationGrid.cpp:224-248), quantised to binary16, classified
    auto idx = [&](int x, int y, int z) {
        x = std::max(x, 0);
        y = std::max(y, 0);
        z = std::max(z, 0);
        return (static_cast<size_t>(z) * ny + y) * nx + x;
    };
    size_t oi = 0, wi = 0;
    ClassKey last = ClassMap::vacuum();
    uint32_t last_id = cm.map.at(last);
    auto classify = [&](size_t c00, float (&es)[3], const float (&ee)[3]) {
        for (; oi < ov.size() && ov[oi].idx == c00; ++oi) es[ov[oi].axis] = ov[oi].sigma;
        float wf[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
        for (; wi < wv.size() && wv[wi].idx == c00; ++wi) wf[wv[wi].slot] = wv[wi].factor;
        ClassKey k;
        for (int c = 0; c < 3; ++c) {
            k[c] = f32_to_f16(std::min(es[c], kPerfectEdgeSigma));
            k[3 + c] = f32_to_f16(ee[c]);
        }
        for (int j = 0; j < 6; ++j) k[6 + j] = f32_to_f16(wf[j]);
        uint32_t id;
        if (k == last) {
            id = last_id;
        } else {
            auto it = cm.map.find(k);
            if (it == cm.map.end()) {
                id = static_cast<uint32_t>(cm.keys.size());
                cm.map.emplace(k, id);
                cm.keys.push_back(k);
            } else {
                id = it->second;
            }
            last = k;
            last_id = id;
        }
        ids[c00] = id;
    };
    if (!sigma && !eps_r) {
        // vacuum volume: every edge average is (0, 1), so only the overridden texels can leave
        // class 0; visiting them in lattice order numbers the classes as the full scan does
        std::fill(ids, ids + static_cast<size_t>(nx) * ny * nz, 0u);
        const float ee[3] = {1.0f, 1.0f, 1.0f};
        while (oi < ov.size() || wi < wv.size()) {
            const size_t a = oi < ov.size() ? ov[oi].idx : SIZE_MAX, b = wi < wv.size() ? wv[wi].idx : SIZE_MAX;
            float es[3] = {0.0f, 0.0f, 0.0f};
            classify(std::min(a, b), es, ee);
        }
        return;
    }
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t c00 = idx(x, y, z);
                const size_t ex[4] = {c00, idx(x, y - 1, z), idx(x, y, z - 1), idx(x, y - 1, z - 1)};
                const size_t ey[4] = {c00, idx(x - 1, y, z), idx(x, y, z - 1), idx(x - 1, y, z - 1)};
                const size_t ez[4] = {c00, idx(x - 1, y, z), idx(x, y - 1, z), idx(x - 1, y - 1, z)};
                auto avg = [](const float *v, const size_t (&c)[4]) {
                    return 0.25f * (v[c[0]] + v[c[1]] + v[c[2]] + v[c[3]]);
                };
                float es[3] = {avg(sigma, ex), avg(sigma, ey), avg(sigma, ez)};
                const float ee[3] = {avg(eps_r, ex), avg(eps_r, ey), avg(eps_r, ez)};
                classify(c00, es, ee);
            }
}

PackedSources pack_sources(const emfd_grid_config &cfg, const emfd_source_desc *sources, int n) {
    PackedSources ps;
    // packing (src/SimulationGrid.cpp:297-318)
    for (int i = 0; i < n && static_cast<int>(ps.kept.size()) < kMaxSources; ++i) {
        const emfd_source_desc &s = sources[i];
        const float len = std::sqrt(s.polarization[0] * s.polarization[0] + s.polarization[1] * s.polarization[1] +
                                    s.polarization[2] * s.polarization[2]);
        if (!s.enabled || !(len > 1e-6f) || !std::isfinite(len)) continue;
        float g[16];
        for (int a = 0; a < 3; ++a) {
            g[a] = s.center[a];
            g[4 + a] = std::clamp(s.half_extent[a], 0.0f, 1.0e5f);
            g[8 + a] = s.polarization[a] / len;
            g[12 + a] = std::max(s.softness[a], 0.25f);
        }
        g[3] = g[7] = g[15] = 0.0f;
        g[11] = s.amplitude;
        ps.packed.insert(ps.packed.end(), g, g + 16);
        ps.kept.push_back(s);
    }
    // per-texel source lists (src/SimulationGrid.cpp:320-349)
    const int nx = cfg.size[0], ny = cfg.size[1], nz = cfg.size[2];
    const int dims[3] = {nx, ny, nz};
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (size_t i = 0; i < ps.kept.size(); ++i) {
        const float *g = &ps.packed[16 * i];
        int lo[3], hi[3];
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::clamp(static_cast<int>(std::floor(g[a] - g[4 + a] - 3.5f * g[12 + a])) - 1, 0, dims[a] - 1);
            hi[a] = std::clamp(static_cast<int>(std::ceil(g[a] + g[4 + a] + 3.5f * g[12 + a])), 0, dims[a] - 1);
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x)
                    pairs.emplace_back(static_cast<uint32_t>((z * ny + y) * nx + x), static_cast<uint32_t>(i));
    }
    std::sort(pairs.begin(), pairs.end());
    for (size_t k = 0; k < pairs.size(); ++k) {
        if (k == 0 || pairs[k].first != pairs[k - 1].first) {
            const uint32_t id = pairs[k].first;
            ps.cells.push_back(static_cast<int>(id % static_cast<uint32_t>(nx)));
            ps.cells.push_back(static_cast<int>((id / static_cast<uint32_t>(nx)) % static_cast<uint32_t>(ny)));
            ps.cells.push_back(static_cast<int>(id / static_cast<uint32_t>(nx * ny)));
            ps.cells.push_back(static_cast<int>(ps.lists.size()));
        }
        ps.lists.push_back(static_cast<int>(pairs[k].second));
    }
    ps.n_cells = static_cast<int>(ps.cells.size() / 4);
    ps.cells.insert(ps.cells.end(), {0, 0, 0, static_cast<int>(ps.lists.size())});
    if (ps.lists.empty()) ps.lists.push_back(0);
    if (ps.packed.empty()) ps.packed.assign(16, 0.0f);
    return ps;
}
}  // namespace emfd

using namespace emfd;

emfd::FieldPtrs emfd_ctx::fields(int which) const {
    FieldPtrs f;
    if (storage == kStorageF32) {
        float *base = field_buf[which].as<float>();
        for (int c = 0; c < 3; ++c) {
            f.e[c] = base + static_cast<size_t>(c) * n;
            f.b[c] = base + static_cast<size_t>(3 + c) * n;
        }
        return f;
    }
    f.storage = storage;
    if (storage == kStorageF64) {
        double *base = field_buf[which].as<double>();
        for (int c = 0; c < 3; ++c) {
            f.e[c] = reinterpret_cast<float *>(base + static_cast<size_t>(c) * n);
            f.b[c] = reinterpret_cast<float *>(base + static_cast<size_t>(3 + c) * n);
        }
        return f;
    }
    // 16-bit planes: the pointers carry 16-bit element addresses
    uint16_t *base = field_buf[which].as<uint16_t>();
    for (int c = 0; c < 3; ++c) {
        f.e[c] = reinterpret_cast<float *>(base + static_cast<size_t>(c) * n);
        f.b[c] = reinterpret_cast<float *>(base + static_cast<size_t>(3 + c) * n);
    }
    f.storage = storage;
    for (int c = 0; c < 6; ++c) {
        f.scale[c] = std::ldexp(1.0f, scale_log2[c]);
        f.inv_scale[c] = std::ldexp(1.0f, -scale_log2[c]);
    }
    return f;
}

emfd::LatticeArgs emfd_ctx::lattice() const {
    LatticeArgs a{};
    a.nx = cfg.size[0];
    a.ny = cfg.size[1];
    a.nz = cfg.size[2];
    a.pml = cfg.pml_thickness;
    a.dt = cfg.courant;
    a.prof_stride = prof_stride;
    a.prof = prof.as<float>();
    a.mat_id = classes > 0 ? mat_id.ptr : nullptr;
    a.id_bytes = id_bytes;
    a.table = table.as<float>();
    a.has_wire = has_wire;
    a.pml_tab = pml_tab.as<float>();
    a.pml_k = 2 * cfg.pml_thickness + 2;
    a.pml_materials = classes > 0 ? pml_materials : 0;
    if (classes > 0 && masks_ok) {
        const int wpr = (cfg.size[0] + 31) / 32;
        a.mask_wpr = wpr;
        a.wire_mask = class_masks.as<uint32_t>();
        a.class_mask = a.wire_mask + static_cast<size_t>(cfg.size[1]) * cfg.size[2] * wpr;
    }
    if (side_active && ntff.enabled && ntff.d_side.ptr) {
        a.side = ntff.side;
        a.side.e_valid = side_e_valid ? 1 : 0;
        a.side.buf = ntff.d_side.as<float>();
    }
    return a;
}

extern "C" {

const char *emfd_last_error(void) { return g_error.c_str(); }

const char *emfd_version(void) {
    return "emfdtd 0.3 gfx950: one-pass z-march and two-pass Yee (FP32 / binary16 / bfloat16 / FP64 storage), "
           "graded absorber or convolutional PML, sources, ports, multi-wavelength far field with FP64 DFTs and MFMA "
           "pattern, flux, probes, coax lines, radar image, volume view, batched stepper, BF16 MFMA MLP";
}

int emfd_create(int device, emfd_ctx **out) {
    if (!out) return fail(EMFD_ERR_INVALID, "emfd_create: out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(EMFD_ERR_HIP, "emfd_create: no HIP device available");
    if (device < 0 || device >= count) return fail(EMFD_ERR_INVALID, "emfd_create: device index out of range");
    EMFD_HIP(hipSetDevice(device));
    // the no-allocation entry points' device array: its address on this device, once
    if (int rc = flux_scratch_init(device)) return rc;
    auto *ctx = new emfd_ctx();
    ctx->device = device;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    *out = ctx;
    return EMFD_OK;
}

void emfd_destroy(emfd_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->field_buf) b.release();
    DevBuf *bufs[] = {&ctx->d_src_descs, &ctx->sigma_cell, &ctx->eps_cell, &ctx->mat_id, &ctx->table, &ctx->prof, &ctx->pml_tab, &ctx->d_packed,
                      &ctx->d_cells, &ctx->d_lists, &ctx->d_src_recs, &ctx->d_src_pairs, &ctx->d_consts, &ctx->ports.d_ports, &ctx->ports.d_phasors,
                      &ctx->ports.d_series, &ctx->ntff.d_phasors, &ctx->ntff.d_side, &ctx->ntff.d_dirs, &ctx->ntff.d_partials,
                      &ctx->ntff.d_out, &ctx->ntff.d_power, &ctx->scratch_partials, &ctx->scratch_misc,
                      &ctx->lines.d_quant, &ctx->lines.d_terms, &ctx->lines.d_weights, &ctx->lines.d_phasors,
                      &ctx->lines.d_now, &ctx->ntff.d_ahi, &ctx->ntff.d_alo, &ctx->ntff.d_absmax, &ctx->fix_b,
                      &ctx->fix_e, &ctx->ports.d_dft, &ctx->ntff.d_dft, &ctx->ntff.d_dft_ph, &ctx->d_consts64,
                      &ctx->d_cpml_coef, &ctx->d_psi, &ctx->ntff.d_samples, &ctx->class_masks, &ctx->march_lid,
                      &ctx->march_tiles, &ctx->march_coef, &ctx->fix_e_rest, &ctx->d_render};
    for (DevBuf *b : bufs) b->release();
    for (hipEvent_t ev : ctx->ev_pool) (void)hipEventDestroy(ev);
    for (int i = 0; i < 2; ++i) {
        if (ctx->h_consts[i]) (void)hipHostFree(ctx->h_consts[i]);
        if (ctx->h_consts_ev[i]) (void)hipEventDestroy(ctx->h_consts_ev[i]);
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int emfd_grid_configure(emfd_ctx *ctx, const emfd_grid_config *requested, emfd_grid_config *applied) {
    if (!ctx || !requested) return fail(EMFD_ERR_INVALID, "emfd_grid_configure: null argument");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));  // buffers may be reallocated below
    ctx->src_recs_dirty = true;
    const emfd_grid_config c = clamp_config(*requested);
    // The reference rejects grids beyond GL_MAX_3D_TEXTURE_SIZE; here the limit is the 32-bit
    // plane index used by the kernels for one component plane.
    const size_t n = static_cast<size_t>(c.size[0]) * c.size[1] * c.size[2];
    if (n > (size_t(1) << 34)) return fail(EMFD_ERR_INVALID, "SimulationGrid: grid too large");

    if (n != ctx->n || !ctx->field_buf[0].ptr) {
        ctx->n = n;
        // the ping-pong buffer serves the fused kernels (FP32 and 16-bit storage)
        for (int k = 0; k < (ctx->storage != kStorageF64 ? 2 : 1); ++k) {
            int rc = ctx->field_buf[k].ensure(ctx->field_bytes());
            if (rc) return rc;
        }
        if (int rc = ctx->sigma_cell.ensure(n * sizeof(float))) return rc;
        if (int rc = ctx->eps_cell.ensure(n * sizeof(float))) return rc;
    }
    ctx->n = n;
    ctx->cfg = c;
    ctx->configured = true;
    ctx->cur = 0;

    ctx->absorber = EMFD_ABSORBER_GRADED;
    int stride = 0;
    const std::vector<float> prof = absorber_profiles(c, &stride);
    ctx->prof_stride = stride;
    if (int rc = ctx->prof.ensure(prof.size() * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpyAsync(ctx->prof.ptr, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice,
                            ctx->stream));
    const size_t kk = static_cast<size_t>(2 * c.pml_thickness + 2);
    if (int rc = ctx->pml_tab.ensure(kk * kk * kk * kPmlStride * sizeof(float))) return rc;
    EMFD_HIP(launch_build_pml_table(ctx->lattice(), ctx->pml_tab.as<float>(), ctx->stream));

    // vacuum everywhere: E = (0,0,0, sigma 0), B = (0,0,0, eps 1) (src/SimulationGrid.cpp:185-192)
    ctx->classes = 0;
    ctx->has_wire = 0;
    ctx->n_fix_b = ctx->n_fix_e = ctx->n_fix_e_rest = 0;
    ctx->march_tables = false;
    ctx->id_bytes = 0;
    ctx->masks_ok = false;
    ctx->pml_materials = 0;  // vacuum everywhere
    EMFD_HIP(hipMemsetAsync(ctx->field_buf[0].ptr, 0, ctx->field_bytes(), ctx->stream));
    EMFD_HIP(hipMemsetAsync(ctx->sigma_cell.ptr, 0, n * sizeof(float), ctx->stream));
    EMFD_HIP(launch_fill(ctx->eps_cell.as<float>(), n, 1.0f, ctx->stream));
    ctx->step_count = 0;
    if (applied) *applied = c;
    return EMFD_OK;
}

int emfd_grid_get_config(emfd_ctx *ctx, emfd_grid_config *out) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    *out = ctx->cfg;
    return EMFD_OK;
}

namespace {
// The texels the one-pass kernel recomputes after its vacuum update (EMFD_KERNEL_MARCH): B of the
// thin-wire classes (a wire factor other than 1), E of every class other than vacuum and of the
// texels whose curl reads a thin-wire texel's B (the texel itself and its +x, +y, +z neighbours).
// Sorted lattice indices.
int build_fix_lists(emfd_ctx *ctx, const std::vector<ClassKey> &keys, const uint32_t *ids) {
    ctx->n_fix_b = ctx->n_fix_e = 0;
    const int nx = ctx->cfg.size[0], ny = ctx->cfg.size[1], nz = ctx->cfg.size[2];
    const size_t n = ctx->n;
    std::vector<char> wire_class(keys.size(), 0);
    for (size_t k = 0; k < keys.size(); ++k)
        for (int f = 6; f < 12; ++f) wire_class[k] |= keys[k][f] != 0x3C00u;  // binary16 1.0
    std::vector<uint32_t> fb, fe, near;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t id = ids[i];
        if (id == 0u) continue;
        fe.push_back(static_cast<uint32_t>(i));
        if (!wire_class[id]) continue;
        fb.push_back(static_cast<uint32_t>(i));
        near.push_back(static_cast<uint32_t>(i));
        const size_t x = i % nx, y = (i / nx) % ny, z = i / (static_cast<size_t>(nx) * ny);
        if (x + 1 < static_cast<size_t>(nx)) near.push_back(static_cast<uint32_t>(i + 1));
        if (y + 1 < static_cast<size_t>(ny)) near.push_back(static_cast<uint32_t>(i + nx));
        if (z + 1 < static_cast<size_t>(nz)) near.push_back(static_cast<uint32_t>(i + static_cast<size_t>(nx) * ny));
    }
    std::sort(near.begin(), near.end());
    near.erase(std::unique(near.begin(), near.end()), near.end());
    auto merge_near = [&near](std::vector<uint32_t> &list) {
        if (near.empty()) return;
        std::vector<uint32_t> merged;
        merged.reserve(list.size() + near.size());
        std::merge(list.begin(), list.end(), near.begin(), near.end(), std::back_inserter(merged));
        merged.erase(std::unique(merged.begin(), merged.end()), merged.end());
        list.swap(merged);
    };

    // The march's material tables (its automatic 8-plane chunks): each texel of a class other than
    // vacuum outside the absorber gets its 1-based position in its tile's class list (up to 255
    // classes per tile); the others (absorber, overflow) stay in the E recomputation list.
    ctx->march_tables = false;
    ctx->march_classes.clear();
    ctx->n_fix_e_rest = 0;
    if (!fe.empty() && nx % 4 == 0 && nz >= kMarchTileZ) {
        const int L = ctx->cfg.pml_thickness;
        const int nseg = (nx + kMarchTileX - 1) / kMarchTileX, nyt = (ny + kMarchTileRows - 1) / kMarchTileRows,
                  nzc = (nz + kMarchTileZ - 1) / kMarchTileZ;
        const size_t n_tiles = static_cast<size_t>(nseg) * nyt * nzc;
        std::vector<std::vector<uint32_t>> tile_cls(n_tiles);
        std::vector<uint8_t> lid(n, 0);
        std::vector<uint32_t> rest;
        size_t last_tile = SIZE_MAX;
        uint32_t last_id = 0;
        uint8_t last_l = 0;
        for (const uint32_t q : fe) {
            const size_t i = q;
            const int x = static_cast<int>(i % nx), y = static_cast<int>((i / nx) % ny),
                      z = static_cast<int>(i / (static_cast<size_t>(nx) * ny));
            const bool absorber = x < L || y < L || z < L || x >= nx - (L + 1) || y >= ny - (L + 1) || z >= nz - (L + 1);
            if (absorber) {
                rest.push_back(q);
                continue;
            }
            const size_t t = static_cast<size_t>(x / kMarchTileX) +
                             nseg * (static_cast<size_t>(y / kMarchTileRows) + static_cast<size_t>(nyt) * (z / kMarchTileZ));
            const uint32_t id = ids[i];
            uint8_t l = 0;
            if (t == last_tile && id == last_id) {
                l = last_l;
            } else {
                std::vector<uint32_t> &cls = tile_cls[t];
                const auto it = std::find(cls.begin(), cls.end(), id);
                if (it != cls.end()) {
                    l = static_cast<uint8_t>(it - cls.begin() + 1);
                } else if (cls.size() < static_cast<size_t>(kMarchMaxLocal)) {
                    cls.push_back(id);
                    l = static_cast<uint8_t>(cls.size());
                }
                last_tile = t;
                last_id = id;
                last_l = l;
            }
            if (l == 0) rest.push_back(q);  // past the tile's 255 classes
            lid[i] = l;
        }
        std::vector<int2> tiles(n_tiles);
        size_t entries = 0;
        for (size_t t = 0; t < n_tiles; ++t) {
            tiles[t] = make_int2(static_cast<int>(entries), static_cast<int>(tile_cls[t].size()));
            entries += tile_cls[t].size();
        }
        if (entries > 0 && entries < 0x7fffffffu) {
            ctx->march_classes.reserve(entries);
            for (const auto &cls : tile_cls) ctx->march_classes.insert(ctx->march_classes.end(), cls.begin(), cls.end());
            if (int rc = ctx->march_lid.ensure(n)) return rc;
            if (int rc = ctx->march_tiles.ensure(n_tiles * sizeof(int2))) return rc;
            EMFD_HIP(hipMemcpy(ctx->march_lid.ptr, lid.data(), n, hipMemcpyHostToDevice));
            EMFD_HIP(hipMemcpy(ctx->march_tiles.ptr, tiles.data(), n_tiles * sizeof(int2), hipMemcpyHostToDevice));
            merge_near(rest);
            if (rest.size() > 0x7fffffffu) return fail(EMFD_ERR_INVALID, "too many material texels");
            if (!rest.empty()) {
                if (int rc = ctx->fix_e_rest.ensure(rest.size() * sizeof(uint32_t))) return rc;
                EMFD_HIP(hipMemcpy(ctx->fix_e_rest.ptr, rest.data(), rest.size() * sizeof(uint32_t),
                                   hipMemcpyHostToDevice));
            }
            ctx->n_fix_e_rest = static_cast<int>(rest.size());
            ctx->march_tables = true;  // once the entries hold the class coefficients (fill_march_coef)
        }
    }
    merge_near(fe);
    if (fb.size() > 0x7fffffffu || fe.size() > 0x7fffffffu) return fail(EMFD_ERR_INVALID, "too many material texels");
    if (!fb.empty()) {
        if (int rc = ctx->fix_b.ensure(fb.size() * sizeof(uint32_t))) return rc;
        EMFD_HIP(hipMemcpy(ctx->fix_b.ptr, fb.data(), fb.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    if (!fe.empty()) {
        if (int rc = ctx->fix_e.ensure(fe.size() * sizeof(uint32_t))) return rc;
        EMFD_HIP(hipMemcpy(ctx->fix_e.ptr, fe.data(), fe.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    ctx->n_fix_b = static_cast<int>(fb.size());
    ctx->n_fix_e = static_cast<int>(fe.size());
    return EMFD_OK;
}

// Entries of the march's material tables: Ca x3, Cb x3 of each listed class, from the class table
// the E kernels read (edge_coefficients: a class other than vacuum outside the absorber).
int fill_march_coef(emfd_ctx *ctx) {
    if (!ctx->march_tables) return EMFD_OK;
    std::vector<float> table(static_cast<size_t>(ctx->classes) * kTableStride);
    EMFD_HIP(hipMemcpy(table.data(), ctx->table.ptr, table.size() * sizeof(float), hipMemcpyDeviceToHost));
    const size_t entries = ctx->march_classes.size();
    std::vector<float> coef(entries * kMarchCoefStride, 0.0f);
    for (size_t k = 0; k < entries; ++k) {
        const float *t = table.data() + static_cast<size_t>(ctx->march_classes[k]) * kTableStride;
        for (int c = 0; c < 3; ++c) {
            coef[k * kMarchCoefStride + c] = t[kCa + c];
            coef[k * kMarchCoefStride + 3 + c] = t[kCb + c];
        }
    }
    if (int rc = ctx->march_coef.ensure(coef.size() * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpy(ctx->march_coef.ptr, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->march_classes.clear();
    ctx->march_classes.shrink_to_fit();
    return EMFD_OK;
}
}  // namespace

int emfd_grid_upload_materials(emfd_ctx *ctx, const float *sigma, const float *eps_r, int64_t n_cells,
                               const emfd_edge_override *edges, int64_t n_edges, const emfd_wire_coefficient *wires,
                               int64_t n_wires) {
    if (!ctx || !sigma || !eps_r) return fail(EMFD_ERR_INVALID, "emfd_grid_upload_materials: null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (n_cells < 0 || static_cast<size_t>(n_cells) != ctx->n)
        return fail(EMFD_ERR_INVALID, "SimulationGrid: material volume size does not match the grid");
    if ((n_edges > 0 && !edges) || (n_wires > 0 && !wires))
        return fail(EMFD_ERR_INVALID, "emfd_grid_upload_materials: null override array");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));  // buffers may be reallocated below
    ctx->src_recs_dirty = true;
    const size_t n = ctx->n;

    // cell materials (alpha channels of the reference's field textures)
    EMFD_HIP(hipMemcpyAsync(ctx->sigma_cell.ptr, sigma, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    EMFD_HIP(hipMemcpyAsync(ctx->eps_cell.ptr, eps_r, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));

    const bool thin = n_wires > 0;  // the reference keys thin-wire mode on a non-empty list
    ClassMap cm;
    std::vector<uint32_t> ids(n);
    classify_cells(ctx->cfg, sigma, eps_r, edges, n_edges, wires, n_wires, cm, ids.data());
    const std::vector<ClassKey> &keys = cm.keys;

    ctx->has_wire = thin ? 1 : 0;
    ctx->masks_ok = false;
    ctx->pml_materials = absorber_has_materials(ctx->cfg, ids.data()) ? 1 : 0;
    if (int rc = build_fix_lists(ctx, keys, ids.data())) return rc;
    if (keys.size() == 1 && !thin) {
        ctx->classes = 0;  // uniform vacuum: kernels skip the class id stream
        ctx->id_bytes = 0;
        ctx->march_tables = false;
    } else {
        ctx->classes = static_cast<int>(keys.size());
        ctx->id_bytes = keys.size() <= 65536 ? 2 : 4;
        if (int rc = ctx->mat_id.ensure(n * static_cast<size_t>(ctx->id_bytes))) return rc;
        if (ctx->id_bytes == 2) {
            std::vector<uint16_t> id16(n);
            for (size_t i = 0; i < n; ++i) id16[i] = static_cast<uint16_t>(ids[i]);
            EMFD_HIP(hipMemcpy(ctx->mat_id.ptr, id16.data(), n * 2, hipMemcpyHostToDevice));
        } else {
            EMFD_HIP(hipMemcpy(ctx->mat_id.ptr, ids.data(), n * 4, hipMemcpyHostToDevice));
        }
        DevBuf d_keys;
        if (int rc = d_keys.ensure(keys.size() * sizeof(ClassKey))) return rc;
        if (int rc = ctx->table.ensure(keys.size() * kTableStride * sizeof(float))) return rc;
        EMFD_HIP(hipMemcpy(d_keys.ptr, keys.data(), keys.size() * sizeof(ClassKey), hipMemcpyHostToDevice));
        EMFD_HIP(launch_build_table(d_keys.as<uint16_t>(), ctx->classes, ctx->cfg.courant, ctx->table.as<float>(),
                                    ctx->stream));
        {
            const int wpr = (ctx->cfg.size[0] + 31) / 32;
            const size_t words = static_cast<size_t>(ctx->cfg.size[1]) * ctx->cfg.size[2] * wpr;
            if (int rc = ctx->class_masks.ensure(2 * words * sizeof(uint32_t))) return rc;
            EMFD_HIP(launch_build_class_masks(ctx->lattice(), ctx->class_masks.as<uint32_t>(),
                                              ctx->class_masks.as<uint32_t>() + words, wpr, ctx->stream));
            ctx->masks_ok = true;
        }
        EMFD_HIP(hipStreamSynchronize(ctx->stream));
        d_keys.release();
        if (int rc = fill_march_coef(ctx)) return rc;
    }
    // uploadMaterials clears E and B and rewinds time (src/SimulationGrid.cpp:216-222,294)
    ctx->cur = 0;
    EMFD_HIP(hipMemsetAsync(ctx->field_buf[0].ptr, 0, ctx->field_bytes(), ctx->stream));
    if (int rc = clear_psi(ctx)) return rc;
    ctx->step_count = 0;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    return EMFD_OK;
}

int emfd_grid_material_classes(emfd_ctx *ctx, int *classes) {
    if (!ctx || !classes) return fail(EMFD_ERR_INVALID, "null argument");
    *classes = ctx->classes;
    return EMFD_OK;
}

int emfd_grid_set_sources(emfd_ctx *ctx, const emfd_source_desc *sources, int n, int *n_active) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (n > 0 && !sources) return fail(EMFD_ERR_INVALID, "null sources");
    EMFD_HIP(hipSetDevice(ctx->device));
    PackedSources ps = pack_sources(ctx->cfg, sources, n);
    ctx->n_cells = ps.n_cells;
    ctx->src_recs_dirty = true;
    std::vector<float> &packed = ps.packed;
    std::vector<int> &cells = ps.cells, &lists = ps.lists;
    const std::vector<emfd_source_desc> &kept = ps.kept;
    if (int rc = ctx->d_packed.ensure(packed.size() * sizeof(float))) return rc;
    if (int rc = ctx->d_src_descs.ensure(std::max<size_t>(kept.size(), 1) * sizeof(emfd_source_desc))) return rc;
    if (int rc = ctx->d_cells.ensure(cells.size() * sizeof(int))) return rc;
    if (int rc = ctx->d_lists.ensure(lists.size() * sizeof(int))) return rc;
    ctx->sources_lists_len = lists.size();
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpy(ctx->d_packed.ptr, packed.data(), packed.size() * sizeof(float), hipMemcpyHostToDevice));
    if (!kept.empty())
        EMFD_HIP(hipMemcpy(ctx->d_src_descs.ptr, kept.data(), kept.size() * sizeof(emfd_source_desc),
                           hipMemcpyHostToDevice));
    EMFD_HIP(hipMemcpy(ctx->d_cells.ptr, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice));
    EMFD_HIP(hipMemcpy(ctx->d_lists.ptr, lists.data(), lists.size() * sizeof(int), hipMemcpyHostToDevice));
    ctx->sources = kept;
    if (n_active) *n_active = static_cast<int>(kept.size());
    return EMFD_OK;
}

int emfd_grid_reset_fields(emfd_ctx *ctx) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    EMFD_HIP(hipSetDevice(ctx->device));
    ctx->cur = 0;
    EMFD_HIP(launch_fill(ctx->field_buf[0].as<float>(), ctx->field_bytes() / 4, 0.0f, ctx->stream));
    if (int rc = clear_psi(ctx)) return rc;
    ctx->step_count = 0;
    return EMFD_OK;
}

int emfd_grid_set_storage(emfd_ctx *ctx, int storage, const int *scale_log2) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (storage != EMFD_STORAGE_F32 && storage != EMFD_STORAGE_F16 && storage != EMFD_STORAGE_BF16 &&
        storage != EMFD_STORAGE_F64)
        return fail(EMFD_ERR_INVALID, "unknown field storage");
    if ((storage == EMFD_STORAGE_F16 || storage == EMFD_STORAGE_BF16) && ctx->cfg.size[0] % 8 != 0)
        return fail(EMFD_ERR_INVALID, "16-bit field storage needs nx % 8 == 0");
    if (storage != EMFD_STORAGE_F32 && ctx->absorber == EMFD_ABSORBER_CPML)
        return fail(EMFD_ERR_STATE, "the convolutional PML runs with FP32 field storage only");
    int k[6] = {0, 0, 0, 0, 0, 0};
    if (scale_log2)
        for (int c = 0; c < 6; ++c) {
            if (scale_log2[c] < -100 || scale_log2[c] > 100) return fail(EMFD_ERR_INVALID, "scale exponent out of range");
            k[c] = scale_log2[c];
        }
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->storage = storage;
    std::copy(k, k + 6, ctx->scale_log2);
    ctx->field_buf[0].release();
    ctx->field_buf[1].release();
    for (int b = 0; b < (storage != EMFD_STORAGE_F64 ? 2 : 1); ++b)
        if (int rc = ctx->field_buf[b].ensure(ctx->field_bytes())) return rc;
    // as uploadMaterials: fields cleared, time rewound
    ctx->cur = 0;
    EMFD_HIP(hipMemsetAsync(ctx->field_buf[0].ptr, 0, ctx->field_bytes(), ctx->stream));
    if (int rc = clear_psi(ctx)) return rc;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->step_count = 0;
    return EMFD_OK;
}

int emfd_grid_set_absorber(emfd_ctx *ctx, int kind, float sigma_max, float grading, float alpha_max) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (kind != EMFD_ABSORBER_GRADED && kind != EMFD_ABSORBER_CPML) return fail(EMFD_ERR_INVALID, "unknown absorber");
    if (kind == EMFD_ABSORBER_CPML && ctx->storage != kStorageF32)
        return fail(EMFD_ERR_STATE, "the convolutional PML runs with FP32 field storage only");
    if (!(sigma_max == sigma_max) || !(grading == grading) || !(alpha_max == alpha_max))
        return fail(EMFD_ERR_INVALID, "NaN absorber parameter");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    // the graded absorber's profiles: the configured ones, or none under the convolutional PML
    emfd_grid_config c = ctx->cfg;
    if (kind == EMFD_ABSORBER_CPML) c.pml_sigma_max = 0.0f;
    int stride = 0;
    const std::vector<float> prof = absorber_profiles(c, &stride);
    EMFD_HIP(hipMemcpy(ctx->prof.ptr, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice));
    ctx->absorber = kind;
    EMFD_HIP(launch_build_pml_table(ctx->lattice(), ctx->pml_tab.as<float>(), ctx->stream));
    ctx->src_recs_dirty = true;  // the sources' Cb in the absorber follow its conductivity
    if (kind == EMFD_ABSORBER_CPML) {
        const int L = ctx->cfg.pml_thickness;
        const double m = grading > 0.0f ? grading : 3.0;
        const double smax = sigma_max > 0.0f ? sigma_max : 0.8 * (m + 1.0);
        const double amax = alpha_max > 0.0f ? alpha_max : 0.0;
        ctx->cpml_sigma_max = static_cast<float>(smax);
        ctx->cpml_grading = static_cast<float>(m);
        ctx->cpml_alpha_max = static_cast<float>(amax);
        // [axis][E position i / B position i + 1/2][i][b, a], as the auxiliary-field recursion reads it
        const int n[3] = {ctx->cfg.size[0], ctx->cfg.size[1], ctx->cfg.size[2]};
        const double dt = static_cast<double>(ctx->cfg.courant);
        std::vector<float> coef(static_cast<size_t>(3) * 2 * stride * 2);
        for (int u = 0; u < 3; ++u)
            for (int kind2 = 0; kind2 < 2; ++kind2)
                for (int i = 0; i < stride; ++i) {
                    float *e = &coef[(static_cast<size_t>(u * 2 + kind2) * stride + i) * 2];
                    e[0] = 1.0f;
                    e[1] = 0.0f;
                    if (i >= n[u] || L <= 0) continue;
                    const double x = i + (kind2 ? 0.5 : 0.0), hi = (n[u] - 1.0) - L;
                    double d = std::max(L - x, x - hi) / L;
                    if (d <= 0.0) continue;
                    d = std::min(d, 1.0);
                    const double sg = smax * std::pow(d, m), al = amax * (1.0 - d);
                    const double b = std::exp(-(sg + al) * dt);
                    e[0] = static_cast<float>(b);
                    e[1] = sg + al > 0.0 ? static_cast<float>(sg / (sg + al) * (b - 1.0)) : 0.0f;
                }
        if (int rc = ctx->d_cpml_coef.ensure(coef.size() * sizeof(float))) return rc;
        EMFD_HIP(hipMemcpy(ctx->d_cpml_coef.ptr, coef.data(), coef.size() * sizeof(float), hipMemcpyHostToDevice));
        const size_t nf = cpml_psi_floats(n[0], n[1], n[2], L);
        if (int rc = ctx->d_psi.ensure(std::max<size_t>(nf, 1) * sizeof(float))) return rc;
    }
    // as uploadMaterials: fields (and auxiliary fields) cleared, time rewound
    ctx->cur = 0;
    EMFD_HIP(launch_fill(ctx->field_buf[0].as<float>(), ctx->field_bytes() / 4, 0.0f, ctx->stream));
    if (int rc = clear_psi(ctx)) return rc;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->step_count = 0;
    return EMFD_OK;
}

int emfd_grid_set_kernel(emfd_ctx *ctx, int variant) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    const int v = variant & 0xff;
    if (v != EMFD_KERNEL_AUTO && v != EMFD_KERNEL_TWO_PASS && v != EMFD_KERNEL_TWO_PASS4 && v != EMFD_KERNEL_MARCH)
        return fail(EMFD_ERR_INVALID, "unknown kernel variant");
    ctx->kernel_variant = v;
    ctx->tile_cfg = variant >> 8;  // the march's z-chunk (0: automatic); the half kernels' cells per thread
    return EMFD_OK;
}

int emfd_grid_kernel_timing(emfd_ctx *ctx, int enable, double *total_ms, int64_t *launches) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    EMFD_HIP(hipSetDevice(ctx->device));
    if (ctx->ev_used > 0) {
        EMFD_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t i = 0; i + 1 < ctx->ev_used; i += 2) {
            float ms = 0.0f;
            EMFD_HIP(hipEventElapsedTime(&ms, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
            ctx->timing_ms += ms;
            ctx->timing_launches += 1;
            if (ctx->timing_samples.size() < (1u << 20)) ctx->timing_samples.push_back(ms);
        }
        ctx->ev_used = 0;
    }
    if (total_ms) *total_ms = ctx->timing_ms;
    if (launches) *launches = ctx->timing_launches;
    if (enable >= 0) {
        if (enable && !ctx->timing_every) {
            ctx->timing_ms = 0.0;
            ctx->timing_launches = 0;
            ctx->timing_samples.clear();
        }
        ctx->timing_every = enable;
    }
    return EMFD_OK;
}

int emfd_grid_kernel_timing_samples(emfd_ctx *ctx, float *out, int cap, int *n) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (cap < 0 || (cap > 0 && !out)) return fail(EMFD_ERR_INVALID, "null or negative sample buffer");
    if (int rc = emfd_grid_kernel_timing(ctx, -1, nullptr, nullptr)) return rc;
    const int have = static_cast<int>(ctx->timing_samples.size());
    const int m = std::min(have, cap);
    for (int i = 0; i < m; ++i) out[i] = ctx->timing_samples[static_cast<size_t>(i)];
    if (n) *n = have;
    return EMFD_OK;
}

namespace {
// EMFD_KERNEL_AUTO runs the one-pass march on FP32 lattices of at least 200^3 cells (8e6; at 256^3
// 0.159 against 0.174 ms per Yee step for the two passes, at 768^3 3.83 against 5.76 ms); below
// that its few long workgroups leave the GPU part idle (128^3: 0.046 against 0.036 ms).
// The march is the automatic choice for FP32 fields from 8e6 texels. With 16-bit storage it halves
// the field bytes but not its instructions, and measured slower than the two 16-bit passes
// (768^3 precision scene 3.44 against 2.92 ms per Yee step, vacuum 2.84 against 2.76 ms): it runs
// there only when asked for. With the convolutional PML the march runs too (its slabs corrected
// after it) when every dimension exceeds the widened slabs, 2 pml + 2 cells.
bool march_auto(const emfd_ctx *ctx) {
    const LatticeArgs lat = ctx->lattice();
    return ctx->storage == kStorageF32 && ctx->n >= 8000000u && yee_march_supported(lat) &&
           (ctx->absorber == EMFD_ABSORBER_GRADED || cpml_march_supported(lat));
}
}  // namespace

int emfd_grid_step(emfd_ctx *ctx, int count) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (count <= 0) return EMFD_OK;
    EMFD_HIP(hipSetDevice(ctx->device));

    // step constants for this call, in one upload: signals at t = (n + 1/2) dt in FP64 cast to
    // float (src/SimulationGrid.cpp:371-387), then the monitors' rotation factors
    const int nsrc = static_cast<int>(ctx->sources.size());
    const int mon = monitors_consts_floats(ctx);
    const int stride = std::max(nsrc + mon, 1);
    const size_t nfl = static_cast<size_t>(stride) * count;
    const int hb = ctx->h_consts_next;
    ctx->h_consts_next ^= 1;
    if (ctx->h_consts_ev[hb]) {
        EMFD_HIP(hipEventSynchronize(ctx->h_consts_ev[hb]));  // its previous copy has run
    } else {
        EMFD_HIP(hipEventCreateWithFlags(&ctx->h_consts_ev[hb], hipEventDisableTiming));
    }
    if (ctx->h_consts_cap[hb] < nfl) {
        if (ctx->h_consts[hb]) EMFD_HIP(hipHostFree(ctx->h_consts[hb]));
        ctx->h_consts[hb] = nullptr;
        EMFD_HIP(hipHostMalloc(reinterpret_cast<void **>(&ctx->h_consts[hb]), nfl * sizeof(float),
                               hipHostMallocDefault));
        ctx->h_consts_cap[hb] = nfl;
    }
    float *hc = ctx->h_consts[hb];
    std::fill(hc, hc + nfl, 0.0f);
    const double dt = static_cast<double>(ctx->cfg.courant);
    for (int k = 0; k < count; ++k) {
        float *row = hc + static_cast<size_t>(k) * stride;
        const double t = (static_cast<double>(ctx->step_count + static_cast<uint64_t>(k)) + 0.5) * dt;
        for (int s = 0; s < nsrc; ++s) row[s] = static_cast<float>(source_signal(ctx->sources[s], t));
        monitors_fill_consts(ctx, row + nsrc, ctx->step_count + static_cast<uint64_t>(k) + 1);
    }
    if (ctx->d_consts.bytes < nfl * sizeof(float)) {
        EMFD_HIP(hipStreamSynchronize(ctx->stream));  // earlier steps may still read the old buffer
        if (int rc = ctx->d_consts.ensure(nfl * sizeof(float))) return rc;
    }
    EMFD_HIP(hipMemcpyAsync(ctx->d_consts.ptr, hc, nfl * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    EMFD_HIP(hipEventRecord(ctx->h_consts_ev[hb], ctx->stream));
    if (int rc = monitors_upload_dft(ctx, count)) return rc;

    {
        // the vectorised FP32 two-pass kernels and the 16-bit kernels copy the far-field x-face
        // columns as they go (the 16-bit ones the stored, rounded values)
        const int v = ctx->kernel_variant;
        const bool tp4 = yee_two_pass4_supported(ctx->lattice()) &&
                         (v == EMFD_KERNEL_AUTO || v == EMFD_KERNEL_TWO_PASS4);
        const bool half = (ctx->storage == kStorageF16 || ctx->storage == kStorageBF16) &&
                          yee_half_supported(ctx->lattice());
        const bool march = (v == EMFD_KERNEL_MARCH && yee_march_supported(ctx->lattice())) ||
                           (v == EMFD_KERNEL_AUTO && march_auto(ctx));
        ctx->side_active = ctx->ntff.enabled && ((ctx->storage == kStorageF32 && (tp4 || march)) || half);
        ctx->side_e_valid = true;
    }
    const LatticeArgs lat = ctx->lattice();
    if (ctx->src_recs_dirty && ctx->n_cells > 0 && ctx->storage != kStorageF64) {
        const size_t pairs = ctx->sources_lists_len;
        if (int rc = ctx->d_src_recs.ensure(static_cast<size_t>(ctx->n_cells) * kSourceRecBytes)) return rc;
        if (int rc = ctx->d_src_pairs.ensure(std::max<size_t>(pairs, 1) * kSourcePairBytes)) return rc;
        EMFD_HIP(launch_build_source_records(lat, ctx->d_cells.as<int>(), ctx->n_cells, ctx->d_lists.as<int>(),
                                             ctx->d_packed.as<float>(), ctx->d_src_recs.ptr, ctx->d_src_pairs.ptr,
                                             ctx->stream));
        ctx->src_recs_dirty = false;
    }
    ctx->ntff.pending = 0;  // a run of deferred far-field samples never spans two calls
    for (int k = 0; k < count; ++k) {
        const float *consts = ctx->d_consts.as<float>() + static_cast<size_t>(k) * stride;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (ctx->timing_every && ctx->step_count % static_cast<uint64_t>(ctx->timing_every) == 0) {
            if (ctx->ev_used + 2 > ctx->ev_pool.size()) {
                hipEvent_t a, b;
                EMFD_HIP(hipEventCreate(&a));
                EMFD_HIP(hipEventCreate(&b));
                ctx->ev_pool.push_back(a);
                ctx->ev_pool.push_back(b);
            }
            ev0 = ctx->ev_pool[ctx->ev_used];
            ev1 = ctx->ev_pool[ctx->ev_used + 1];
            ctx->ev_used += 2;
            EMFD_HIP(hipEventRecord(ev0, ctx->stream));
        }
        int variant = ctx->kernel_variant;
        if (ctx->storage != kStorageF32) {
            // 16-bit storage: the two passes, or the march when asked for; FP64: its own two passes
            const bool march16 = ctx->storage != kStorageF64 &&
                                 (variant == EMFD_KERNEL_MARCH || (variant == EMFD_KERNEL_AUTO && march_auto(ctx)));
            if (variant != EMFD_KERNEL_AUTO && !march16)
                return fail(EMFD_ERR_INVALID, "16-bit and FP64 field storage run the automatic kernel or the march");
            if (ctx->storage == kStorageF64) {
                EMFD_HIP(launch_yee_f64(lat, ctx->fields(ctx->cur), ctx->stream));
                variant = -1;
            } else if (march16) {
                variant = EMFD_KERNEL_MARCH;
            } else {
                EMFD_HIP(launch_yee_half(lat, ctx->fields(ctx->cur), ctx->tile_cfg & 0xf, ctx->stream));
                variant = -1;
            }
        } else if (variant == EMFD_KERNEL_AUTO)
            variant = march_auto(ctx) ? EMFD_KERNEL_MARCH
                                      : yee_two_pass4_supported(lat) ? EMFD_KERNEL_TWO_PASS4 : EMFD_KERNEL_TWO_PASS;
        if (variant == EMFD_KERNEL_TWO_PASS4 && !yee_two_pass4_supported(lat))
            return fail(EMFD_ERR_INVALID, "the vectorised two-pass kernel needs nx % 4 == 0");
        if (variant == EMFD_KERNEL_MARCH && !yee_march_supported(lat))
            return fail(EMFD_ERR_INVALID, "the one-pass march kernel needs nx % 4 == 0 and planes of at most 2 GB");
        if (variant == EMFD_KERNEL_MARCH && ctx->absorber == EMFD_ABSORBER_CPML && !cpml_march_supported(lat))
            return fail(EMFD_ERR_INVALID, "the convolutional PML with the march needs every dimension above 2 pml + 2");
        if (variant == -1) {
            // launched above
        } else if (ctx->absorber == EMFD_ABSORBER_CPML && variant != EMFD_KERNEL_MARCH) {
            // Faraday pass, its auxiliary fields, Ampere pass, its auxiliary fields
            const FieldPtrs f = ctx->fields(ctx->cur);
            const float *coef = ctx->d_cpml_coef.as<float>();
            float *psi = ctx->d_psi.as<float>();
            for (int ph = 0; ph < 2; ++ph) {
                if (variant == EMFD_KERNEL_TWO_PASS4)
                    EMFD_HIP(launch_yee_two_pass4(lat, f, ctx->stream, 1 << ph));
                else
                    EMFD_HIP(launch_yee_two_pass(lat, f, ctx->stream, 1 << ph));
                EMFD_HIP(launch_cpml(lat, f, coef, psi, ph == 1, ctx->stream));
            }
        } else if (variant == EMFD_KERNEL_TWO_PASS) {
            EMFD_HIP(launch_yee_two_pass(lat, ctx->fields(ctx->cur), ctx->stream));
        } else if (variant == EMFD_KERNEL_TWO_PASS4) {
            EMFD_HIP(launch_yee_two_pass4(lat, ctx->fields(ctx->cur), ctx->stream));
        } else {
            // bits 8.. of the variant: the march's z-chunk, or under the automatic choice with 16-bit
            // storage the two-pass kernels' cells per thread (the march then picks its z-chunk)
            const int zc = ctx->storage != kStorageF32 && ctx->kernel_variant == EMFD_KERNEL_AUTO ? 0 : ctx->tile_cfg;
            const MarchCpml cpml{ctx->d_cpml_coef.as<float>(), ctx->d_psi.as<float>()};
            EMFD_HIP(launch_yee_march(lat, ctx->fields(ctx->cur), ctx->fields(ctx->cur ^ 1), zc,
                                      ctx->fix_b.as<uint32_t>(), ctx->n_fix_b, ctx->fix_e.as<uint32_t>(), ctx->n_fix_e,
                                      ctx->march_mat(), ctx->fix_e_rest.as<uint32_t>(), ctx->n_fix_e_rest,
                                      ctx->absorber == EMFD_ABSORBER_CPML ? &cpml : nullptr, ctx->stream));
            ctx->cur ^= 1;
        }
        if (ev1) EMFD_HIP(hipEventRecord(ev1, ctx->stream));
        if (ctx->n_cells > 0 && nsrc > 0) {
            if (ctx->storage == kStorageF64) {
                const double t = (static_cast<double>(ctx->step_count) + 0.5) * static_cast<double>(ctx->cfg.courant);
                EMFD_HIP(launch_sources_f64(lat, ctx->fields(ctx->cur), ctx->d_cells.as<int>(), ctx->n_cells,
                                            ctx->d_lists.as<int>(), ctx->d_packed.as<float>(),
                                            ctx->d_src_descs.as<emfd_source_desc>(), t, ctx->stream));
            } else {
                EMFD_HIP(launch_apply_source_records(lat, ctx->fields(ctx->cur), ctx->d_src_recs.ptr,
                                                     ctx->d_src_pairs.ptr, ctx->n_cells, ctx->d_packed.as<float>(),
                                                     consts, ctx->stream));
            }
        }
        ++ctx->step_count;  // observers see the time of the step that just completed
        if (int rc = monitors_after_step(ctx, consts + nsrc, stride, k, count)) return rc;
        if (ctx->ev_used >= 4096) {
            double ms;
            int64_t l;
            if (int rc = emfd_grid_kernel_timing(ctx, -1, &ms, &l)) return rc;
        }
    }
    return EMFD_OK;
}

int emfd_grid_synchronize(emfd_ctx *ctx) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    return EMFD_OK;
}

int emfd_grid_state(emfd_ctx *ctx, double *time, uint64_t *step_count, float *dt) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (time) *time = ctx->time();
    if (step_count) *step_count = ctx->step_count;
    if (dt) *dt = ctx->cfg.courant;
    return EMFD_OK;
}

namespace {
// Component planes of the current fields as FP32 values on the host (3 planes of n for E or B).
int read_planes(emfd_ctx *ctx, int which, std::vector<float> &out) {
    const size_t n = ctx->n;
    out.resize(3 * n);
    if (ctx->storage == kStorageF32) {
        EMFD_HIP(hipMemcpy(out.data(), ctx->field_buf[ctx->cur].as<float>() + static_cast<size_t>(3 * which) * n,
                           3 * n * sizeof(float), hipMemcpyDeviceToHost));
        return EMFD_OK;
    }
    if (ctx->storage == kStorageF64) {
        std::vector<double> d(3 * n);
        EMFD_HIP(hipMemcpy(d.data(), ctx->field_buf[ctx->cur].as<double>() + static_cast<size_t>(3 * which) * n,
                           3 * n * sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < 3 * n; ++i) out[i] = static_cast<float>(d[i]);
        return EMFD_OK;
    }
    std::vector<uint16_t> raw(3 * n);
    EMFD_HIP(hipMemcpy(raw.data(), ctx->field_buf[ctx->cur].as<uint16_t>() + static_cast<size_t>(3 * which) * n,
                       3 * n * 2, hipMemcpyDeviceToHost));
    for (int c = 0; c < 3; ++c) {
        const float s = std::ldexp(1.0f, ctx->scale_log2[3 * which + c]);
        for (size_t i = 0; i < n; ++i) {
            const uint16_t h = raw[static_cast<size_t>(c) * n + i];
            float v;
            if (ctx->storage == kStorageF16) {
                _Float16 x;
                std::memcpy(&x, &h, 2);
                v = static_cast<float>(x);
            } else {
                const uint32_t u = static_cast<uint32_t>(h) << 16;
                std::memcpy(&v, &u, 4);
            }
            out[static_cast<size_t>(c) * n + i] = v * s;
        }
    }
    return EMFD_OK;
}

uint16_t f32_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

int write_planes(emfd_ctx *ctx, int which, const std::vector<float> &in) {
    const size_t n = ctx->n;
    if (ctx->storage == kStorageF32) {
        EMFD_HIP(hipMemcpy(ctx->field_buf[ctx->cur].as<float>() + static_cast<size_t>(3 * which) * n, in.data(),
                           3 * n * sizeof(float), hipMemcpyHostToDevice));
        return EMFD_OK;
    }
    if (ctx->storage == kStorageF64) {
        std::vector<double> d(in.begin(), in.end());
        EMFD_HIP(hipMemcpy(ctx->field_buf[ctx->cur].as<double>() + static_cast<size_t>(3 * which) * n, d.data(),
                           3 * n * sizeof(double), hipMemcpyHostToDevice));
        return EMFD_OK;
    }
    std::vector<uint16_t> raw(3 * n);
    for (int c = 0; c < 3; ++c) {
        const float inv = std::ldexp(1.0f, -ctx->scale_log2[3 * which + c]);
        for (size_t i = 0; i < n; ++i) {
            const float v = in[static_cast<size_t>(c) * n + i] * inv;
            raw[static_cast<size_t>(c) * n + i] = ctx->storage == kStorageF16 ? f32_to_f16(v) : f32_to_bf16(v);
        }
    }
    EMFD_HIP(hipMemcpy(ctx->field_buf[ctx->cur].as<uint16_t>() + static_cast<size_t>(3 * which) * n, raw.data(),
                       3 * n * 2, hipMemcpyHostToDevice));
    return EMFD_OK;
}
}  // namespace

int emfd_grid_download_fields(emfd_ctx *ctx, float *E, float *B, int layout) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (layout != EMFD_LAYOUT_TEXEL && layout != EMFD_LAYOUT_SOA) return fail(EMFD_ERR_INVALID, "unknown layout");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    const size_t n = ctx->n;
    if (layout == EMFD_LAYOUT_SOA && ctx->storage == kStorageF32) {
        const float *base = ctx->field_buf[ctx->cur].as<float>();
        if (E) EMFD_HIP(hipMemcpy(E, base, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
        if (B) EMFD_HIP(hipMemcpy(B, base + 3 * n, 3 * n * sizeof(float), hipMemcpyDeviceToHost));
        return EMFD_OK;
    }
    std::vector<float> plane, alpha(n);
    for (int f = 0; f < 2; ++f) {
        float *dst = f == 0 ? E : B;
        if (!dst) continue;
        if (int rc = read_planes(ctx, f, plane)) return rc;
        if (layout == EMFD_LAYOUT_SOA) {
            std::copy(plane.begin(), plane.end(), dst);
            continue;
        }
        EMFD_HIP(hipMemcpy(alpha.data(), f == 0 ? ctx->sigma_cell.ptr : ctx->eps_cell.ptr, n * sizeof(float),
                           hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; ++i) {
            dst[4 * i + 0] = plane[i];
            dst[4 * i + 1] = plane[n + i];
            dst[4 * i + 2] = plane[2 * n + i];
            dst[4 * i + 3] = alpha[i];
        }
    }
    return EMFD_OK;
}

int emfd_grid_download_fields_f64(emfd_ctx *ctx, double *E, double *B) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    const size_t n = ctx->n;
    for (int f = 0; f < 2; ++f) {
        double *dst = f == 0 ? E : B;
        if (!dst) continue;
        if (ctx->storage == kStorageF64) {
            EMFD_HIP(hipMemcpy(dst, ctx->field_buf[ctx->cur].as<double>() + static_cast<size_t>(3 * f) * n,
                               3 * n * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> plane;
            if (int rc = read_planes(ctx, f, plane)) return rc;
            for (size_t i = 0; i < 3 * n; ++i) dst[i] = plane[i];
        }
    }
    return EMFD_OK;
}

int emfd_grid_upload_fields(emfd_ctx *ctx, const float *E, const float *B, int layout) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (layout != EMFD_LAYOUT_TEXEL && layout != EMFD_LAYOUT_SOA) return fail(EMFD_ERR_INVALID, "unknown layout");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    const size_t n = ctx->n;
    for (int f = 0; f < 2; ++f) {
        const float *srcp = f == 0 ? E : B;
        if (!srcp) continue;
        std::vector<float> plane(3 * n);
        if (layout == EMFD_LAYOUT_SOA) {
            std::copy(srcp, srcp + 3 * n, plane.begin());
        } else {
            for (size_t i = 0; i < n; ++i)
                for (int c = 0; c < 3; ++c) plane[static_cast<size_t>(c) * n + i] = srcp[4 * i + c];
        }
        if (int rc = write_planes(ctx, f, plane)) return rc;
    }
    return EMFD_OK;
}

int emfd_copy_probe(emfd_ctx *ctx, int64_t bytes, int iterations, double *avg_ms) {
    if (!ctx || bytes < 16 || iterations < 1) return fail(EMFD_ERR_INVALID, "emfd_copy_probe: bad arguments");
    EMFD_HIP(hipSetDevice(ctx->device));
    DevBuf a, b;
    if (int rc = a.ensure(static_cast<size_t>(bytes))) return rc;
    if (int rc = b.ensure(static_cast<size_t>(bytes))) return rc;
    EMFD_HIP(launch_fill(a.as<float>(), static_cast<size_t>(bytes) / 4, 1.0f, ctx->stream));
    EMFD_HIP(launch_copy4(a.ptr, b.ptr, static_cast<size_t>(bytes), ctx->stream));  // warm-up
    hipEvent_t e0, e1;
    EMFD_HIP(hipEventCreate(&e0));
    EMFD_HIP(hipEventCreate(&e1));
    EMFD_HIP(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iterations; ++i) EMFD_HIP(launch_copy4(a.ptr, b.ptr, static_cast<size_t>(bytes), ctx->stream));
    EMFD_HIP(hipEventRecord(e1, ctx->stream));
    EMFD_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    EMFD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (avg_ms) *avg_ms = ms / iterations;
    return EMFD_OK;
}

int emfd_stream_probe(emfd_ctx *ctx, int64_t plane_bytes, int planes, int iterations, double *avg_ms) {
    if (!ctx || plane_bytes < (1 << 23) || iterations < 1 || (planes != 1 && planes != 2 && planes != 3 && planes != 6))
        return fail(EMFD_ERR_INVALID, "emfd_stream_probe: bad arguments");
    EMFD_HIP(hipSetDevice(ctx->device));
    const size_t unit = static_cast<size_t>(256) * 2 * 4096;  // float4 per block-round of the probe kernel
    const size_t plane_f4 = static_cast<size_t>(plane_bytes) / 16 / unit * unit;
    DevBuf a, b;
    if (int rc = a.ensure(plane_f4 * 16 * planes)) return rc;
    if (int rc = b.ensure(plane_f4 * 16 * planes)) return rc;
    EMFD_HIP(launch_fill(a.as<float>(), plane_f4 * 4 * planes, 1.0f, ctx->stream));
    EMFD_HIP(launch_stream_probe(a.ptr, b.ptr, plane_f4, planes, ctx->stream));  // warm-up
    hipEvent_t e0, e1;
    EMFD_HIP(hipEventCreate(&e0));
    EMFD_HIP(hipEventCreate(&e1));
    EMFD_HIP(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iterations; ++i) EMFD_HIP(launch_stream_probe(a.ptr, b.ptr, plane_f4, planes, ctx->stream));
    EMFD_HIP(hipEventRecord(e1, ctx->stream));
    EMFD_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    EMFD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (avg_ms) *avg_ms = ms / iterations;
    return EMFD_OK;
}

int64_t emfd_grid_memory_bytes(emfd_ctx *ctx) {
    if (!ctx) return -1;
    int64_t total = 0;
    const DevBuf *bufs[] = {&ctx->field_buf[0], &ctx->field_buf[1], &ctx->sigma_cell, &ctx->eps_cell, &ctx->mat_id,
                            &ctx->table, &ctx->prof, &ctx->d_packed, &ctx->d_cells, &ctx->d_lists, &ctx->d_consts};
    for (const DevBuf *b : bufs) total += static_cast<int64_t>(b->bytes);
    return total;
}

}  // extern "C"
