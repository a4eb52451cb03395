// libemfdtd.so host side of the batched stepper (emfd_batch_*). Each scene is set up with the
// same host functions as a single grid (configure clamps, absorber profiles, material classes,
// source packing, port packing) and its per-step constants come from the same FP64 formulas, so
// the device sees exactly the values a single-grid context would. Scenes keep their own field,
// class-id and table buffers; the launch-level state is a scene descriptor array, a flat tile
// list for the Yee passes, flat source-texel and port lists, and a per-step constant table.
#include "context.h"
#include "host_common.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>

namespace emfd {
namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr size_t kSeriesFloats = size_t(1) << 23;  // PortMonitor's series cap (src/Ports.cpp:156)
constexpr int kTileThreads = 256;

uint32_t magic_of(uint32_t d) {
    return static_cast<uint32_t>(((uint64_t(1) << 32) + d - 1) / d);
}

struct SceneHost {
    emfd_grid_config cfg{};
    size_t n = 0;
    DevBuf fields;  // E x y z, B x y z planes
    DevBuf ids, table, prof;
    DevBuf masks;  // thin-wire mask then class-nonzero mask (LatticeArgs), when there are classes
    int prof_stride = 0;
    int id_bytes = 0, classes = 0, has_wire = 0;
    int pml_materials = 0;  // LatticeArgs::pml_materials
    int pml_table = 0;  // index into emfd_batch::pml_tabs
    std::vector<emfd_source_desc> sources;
    int n_src_cells = 0;
    DevBuf d_packed, d_cells, d_lists;
    std::vector<emfd_port_setup> ports;
    float averaging = 4.0f, wavelength = 16.0f, blend = 0.0f;
    bool has_sweep = false;
    emfd_sweep_setup sweep{};
    DevBuf d_ports, d_phasors, d_series;
    int sig_offset = 0, rot_offset = 0;
    // sweep results
    bool evaluated = false;
    std::vector<emfd_sweep_point> points;
    std::vector<emfd_port_sweep> summary;
    double residual = 1.0;

    void release() {
        for (DevBuf *b : {&fields, &ids, &table, &prof, &masks, &d_packed, &d_cells, &d_lists, &d_ports, &d_phasors,
                          &d_series})
            b->release();
    }
    LatticeArgs lattice(const float *pml_tab) const {
        LatticeArgs a{};
        a.nx = cfg.size[0];
        a.ny = cfg.size[1];
        a.nz = cfg.size[2];
        a.pml = cfg.pml_thickness;
        a.dt = cfg.courant;
        a.prof_stride = prof_stride;
        a.prof = prof.as<float>();
        a.mat_id = classes > 0 ? ids.ptr : nullptr;
        a.id_bytes = id_bytes;
        a.table = table.as<float>();
        a.has_wire = has_wire;
        a.pml_tab = pml_tab;
        a.pml_k = 2 * cfg.pml_thickness + 2;
        a.pml_materials = classes > 0 ? pml_materials : 0;
        if (classes > 0 && masks.ptr) {
            a.mask_wpr = (cfg.size[0] + 31) / 32;
            a.wire_mask = masks.as<uint32_t>();
            a.class_mask = a.wire_mask + static_cast<size_t>(cfg.size[1]) * cfg.size[2] * a.mask_wpr;
        }
        return a;
    }
    FieldPtrs field_ptrs() const {
        FieldPtrs f;
        float *base = fields.as<float>();
        for (int c = 0; c < 3; ++c) {
            f.e[c] = base + static_cast<size_t>(c) * n;
            f.b[c] = base + static_cast<size_t>(3 + c) * n;
        }
        return f;
    }
};

struct PmlKey {
    int pml;
    float courant, sigma_max, grading;
    bool operator==(const PmlKey &o) const {
        return pml == o.pml && std::memcmp(&courant, &o.courant, 4) == 0 &&
               std::memcmp(&sigma_max, &o.sigma_max, 4) == 0 && std::memcmp(&grading, &o.grading, 4) == 0;
    }
};

}  // namespace
}  // namespace emfd

using namespace emfd;

struct emfd_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 256;
    std::vector<std::unique_ptr<SceneHost>> scenes;
    std::vector<PmlKey> pml_keys;
    std::vector<DevBuf> pml_tabs;
    bool finalized = false;
    int64_t max_steps = 0;
    uint64_t step_count = 0;
    int64_t total_cells = 0;
    int64_t stride = 0;  // floats per step row of the constant table
    DevBuf d_scenes, d_tiles, d_src_list, d_port_list, d_consts, d_step;
    int n_tiles = 0, n_src = 0, n_port_items = 0;
    emfd_batch_launch launch{1, 0, 8};
    // graph
    int graph_steps = 0;
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_launches;
    double timed_ms = 0.0;
    int64_t timed_launches = 0;
    // sweep scratch
    DevBuf d_jobs, d_sweep_steps, d_sums, d_residual;

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
    int build_tiles();
    int enqueue_steps(int count);
    BatchLaunchArgs launch_args(const emfd_batch_launch &l) const;
    int collect_timing();
};

int emfd_batch::build_tiles() {
    const int Q = launch.quads_per_thread;
    const uint32_t per = static_cast<uint32_t>(kTileThreads * Q);
    std::vector<int2> tiles;
    for (size_t s = 0; s < scenes.size(); ++s) {
        const uint32_t nq = static_cast<uint32_t>(scenes[s]->n / 4);
        for (uint32_t q0 = 0; q0 < nq; q0 += per) tiles.push_back(make_int2(static_cast<int>(s), static_cast<int>(q0)));
    }
    n_tiles = static_cast<int>(tiles.size());
    if (int rc = d_tiles.ensure(std::max<size_t>(tiles.size(), 1) * sizeof(int2))) return rc;
    if (!tiles.empty())
        EMFD_HIP(hipMemcpy(d_tiles.ptr, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
    return EMFD_OK;
}

// Launches `count` steps (B, E, sources, ports per step) then advances the device step counter.
// Everything here is graph-capturable: kernel launches only.
BatchLaunchArgs emfd_batch::launch_args(const emfd_batch_launch &l) const {
    BatchLaunchArgs la{};
    la.scenes = d_scenes.as<BatchScene>();
    la.tiles = d_tiles.as<int2>();
    la.n_tiles = n_tiles;
    la.slices = l.xcd_slices;
    la.quads = l.quads_per_thread;
    la.pml_materials = 0;
    for (const auto &sc : scenes) la.pml_materials |= sc->classes > 0 ? sc->pml_materials : 0;
    return la;
}

int emfd_batch::enqueue_steps(int count) {
    const BatchLaunchArgs la = launch_args(launch);
    const float *consts = d_consts.as<float>();
    int64_t *ds = d_step.as<int64_t>();
    for (int k = 0; k < count; ++k) {
        EMFD_HIP(launch_batch_yee(la, launch.persistent_blocks, stream));
        EMFD_HIP(launch_batch_sources(d_scenes.as<BatchScene>(), d_src_list.as<int2>(), n_src, consts, stride, ds, k,
                                      stream));
        EMFD_HIP(launch_batch_ports(d_scenes.as<BatchScene>(), d_port_list.as<int2>(), n_port_items, consts, stride,
                                    ds, k, stream));
    }
    EMFD_HIP(launch_batch_advance(ds, count, stream));
    return EMFD_OK;
}

int emfd_batch::collect_timing() {
    EMFD_HIP(hipStreamSynchronize(stream));
    for (size_t i = 0; i + 1 < ev_used; i += 2) {
        float ms = 0.0f;
        EMFD_HIP(hipEventElapsedTime(&ms, ev_pool[i], ev_pool[i + 1]));
        timed_ms += ms;
        timed_launches += ev_launches[i / 2];
    }
    ev_used = 0;
    ev_launches.clear();
    return EMFD_OK;
}

extern "C" {

int emfd_batch_create(int device, emfd_batch **out) {
    if (!out) return fail(EMFD_ERR_INVALID, "null output pointer");
    *out = nullptr;
    int count = 0;
    EMFD_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(EMFD_ERR_INVALID, "no such HIP device");
    EMFD_HIP(hipSetDevice(device));
    std::unique_ptr<emfd_batch> b(new emfd_batch());
    b->device = device;
    EMFD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        b->cus = cus;
    if (int rc = b->d_step.ensure(sizeof(int64_t))) return rc;
    EMFD_HIP(hipMemset(b->d_step.ptr, 0, sizeof(int64_t)));
    *out = b.release();
    return EMFD_OK;
}

int emfd_batch_destroy(emfd_batch *b) {
    if (!b) return EMFD_OK;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->drop_graph();
    for (auto &s : b->scenes) s->release();
    for (auto &t : b->pml_tabs) t.release();
    for (DevBuf *d : {&b->d_scenes, &b->d_tiles, &b->d_src_list, &b->d_port_list, &b->d_consts, &b->d_step, &b->d_jobs,
                      &b->d_sweep_steps, &b->d_sums, &b->d_residual})
        d->release();
    for (hipEvent_t e : b->ev_pool) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return EMFD_OK;
}

int emfd_batch_add_scene(emfd_batch *b, const emfd_batch_scene *sc, int *index, emfd_grid_config *applied) {
    if (!b || !sc) return fail(EMFD_ERR_INVALID, "null argument");
    if (b->finalized) return fail(EMFD_ERR_STATE, "batch already finalized");
    if ((sc->sigma == nullptr) != (sc->eps_r == nullptr))
        return fail(EMFD_ERR_INVALID, "sigma and eps_r must both be given or both be NULL");
    if ((sc->n_edges > 0 && !sc->edges) || (sc->n_wires > 0 && !sc->wires) || (sc->n_sources > 0 && !sc->sources) ||
        (sc->n_ports > 0 && !sc->ports))
        return fail(EMFD_ERR_INVALID, "null array with a non-zero count");
    if (sc->n_ports > EMFD_BATCH_MAX_PORTS) return fail(EMFD_ERR_INVALID, "too many ports for a batched scene");
    EMFD_HIP(hipSetDevice(b->device));
    std::unique_ptr<SceneHost> s(new SceneHost());
    s->cfg = clamp_config(sc->config);
    const int nx = s->cfg.size[0], ny = s->cfg.size[1], nz = s->cfg.size[2];
    if (nx % 4 != 0) return fail(EMFD_ERR_INVALID, "batched scenes need nx % 4 == 0");
    s->n = static_cast<size_t>(nx) * ny * nz;
    if (sc->sigma && (sc->n_cells < 0 || static_cast<size_t>(sc->n_cells) != s->n))
        return fail(EMFD_ERR_INVALID, "SimulationGrid: material volume size does not match the grid");
    const uint64_t nq = s->n / 4, qx = static_cast<uint64_t>(nx / 4);
    if (nq * qx >= (uint64_t(1) << 32) || static_cast<uint64_t>(ny) * nz * ny >= (uint64_t(1) << 32))
        return fail(EMFD_ERR_INVALID, "batched scene too large for 32-bit quad indexing");

    // fields, absorber profiles, PML coefficient table (shared by equal absorber settings)
    if (int rc = s->fields.ensure(s->n * 6 * sizeof(float))) return rc;
    EMFD_HIP(hipMemsetAsync(s->fields.ptr, 0, s->n * 6 * sizeof(float), b->stream));
    int stride = 0;
    const std::vector<float> prof = absorber_profiles(s->cfg, &stride);
    s->prof_stride = stride;
    if (int rc = s->prof.ensure(prof.size() * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpy(s->prof.ptr, prof.data(), prof.size() * sizeof(float), hipMemcpyHostToDevice));
    const PmlKey key{s->cfg.pml_thickness, s->cfg.courant, s->cfg.pml_sigma_max, s->cfg.pml_grading};
    auto it = std::find(b->pml_keys.begin(), b->pml_keys.end(), key);
    if (it == b->pml_keys.end()) {
        const int kk = 2 * s->cfg.pml_thickness + 2;
        DevBuf tab;
        if (int rc = tab.ensure(static_cast<size_t>(kk) * kk * kk * kPmlStride * sizeof(float))) return rc;
        EMFD_HIP(launch_build_pml_table(s->lattice(nullptr), tab.as<float>(), b->stream));
        b->pml_keys.push_back(key);
        b->pml_tabs.push_back(tab);
        s->pml_table = static_cast<int>(b->pml_tabs.size()) - 1;
    } else {
        s->pml_table = static_cast<int>(it - b->pml_keys.begin());
    }

    // material classes (uploadMaterials)
    const bool thin = sc->n_wires > 0;
    ClassMap cm;
    std::vector<uint32_t> ids(s->n);
    classify_cells(s->cfg, sc->sigma, sc->eps_r, sc->edges, sc->n_edges, sc->wires, sc->n_wires, cm, ids.data());
    s->has_wire = thin ? 1 : 0;
    s->pml_materials = absorber_has_materials(s->cfg, ids.data()) ? 1 : 0;
    if (!(cm.keys.size() == 1 && !thin)) {
        s->classes = static_cast<int>(cm.keys.size());
        s->id_bytes = cm.keys.size() <= 65536 ? 2 : 4;
        if (int rc = s->ids.ensure(s->n * static_cast<size_t>(s->id_bytes))) return rc;
        if (s->id_bytes == 2) {
            std::vector<uint16_t> id16(s->n);
            for (size_t i = 0; i < s->n; ++i) id16[i] = static_cast<uint16_t>(ids[i]);
            EMFD_HIP(hipMemcpy(s->ids.ptr, id16.data(), s->n * 2, hipMemcpyHostToDevice));
        } else {
            EMFD_HIP(hipMemcpy(s->ids.ptr, ids.data(), s->n * 4, hipMemcpyHostToDevice));
        }
        DevBuf d_keys;
        if (int rc = d_keys.ensure(cm.keys.size() * sizeof(ClassKey))) return rc;
        if (int rc = s->table.ensure(cm.keys.size() * kTableStride * sizeof(float))) return rc;
        EMFD_HIP(hipMemcpy(d_keys.ptr, cm.keys.data(), cm.keys.size() * sizeof(ClassKey), hipMemcpyHostToDevice));
        EMFD_HIP(launch_build_table(d_keys.as<uint16_t>(), s->classes, s->cfg.courant, s->table.as<float>(), b->stream));
        const int wpr = (s->cfg.size[0] + 31) / 32;
        const size_t words = static_cast<size_t>(s->cfg.size[1]) * s->cfg.size[2] * wpr;
        if (int rc = s->masks.ensure(2 * words * sizeof(uint32_t))) return rc;
        EMFD_HIP(launch_build_class_masks(s->lattice(nullptr), s->masks.as<uint32_t>(), s->masks.as<uint32_t>() + words,
                                          wpr, b->stream));
        EMFD_HIP(hipStreamSynchronize(b->stream));
        d_keys.release();
    }

    // sources (setSources)
    PackedSources ps = pack_sources(s->cfg, sc->sources, std::max(sc->n_sources, 0));
    s->sources = ps.kept;
    s->n_src_cells = ps.n_cells;
    if (int rc = s->d_packed.ensure(ps.packed.size() * sizeof(float))) return rc;
    if (int rc = s->d_cells.ensure(ps.cells.size() * sizeof(int))) return rc;
    if (int rc = s->d_lists.ensure(ps.lists.size() * sizeof(int))) return rc;
    EMFD_HIP(hipMemcpy(s->d_packed.ptr, ps.packed.data(), ps.packed.size() * sizeof(float), hipMemcpyHostToDevice));
    EMFD_HIP(hipMemcpy(s->d_cells.ptr, ps.cells.data(), ps.cells.size() * sizeof(int), hipMemcpyHostToDevice));
    EMFD_HIP(hipMemcpy(s->d_lists.ptr, ps.lists.data(), ps.lists.size() * sizeof(int), hipMemcpyHostToDevice));

    // ports (PortMonitor::configure, src/Ports.cpp:187-198)
    s->ports.assign(sc->ports, sc->ports + std::max(sc->n_ports, 0));
    s->averaging = std::max(sc->averaging_periods, 0.5f);
    s->wavelength = s->ports.empty() ? 16.0f : std::max(s->ports.front().wavelength, 2.0f);
    s->blend = port_blend(s->averaging, s->wavelength, s->cfg.courant);
    s->has_sweep = sc->sweep != nullptr;
    if (sc->sweep) s->sweep = *sc->sweep;
    if (!s->ports.empty()) {
        std::vector<PortDev> packed;
        for (const emfd_port_setup &p : s->ports) {
            PortDev d{};
            for (int a = 0; a < 3; ++a) d.texel[a] = p.texel[a];
            d.axis = std::clamp(p.axis, 0, 2);
            d.edges = p.edges;
            d.sign = p.sign >= 0.0f ? 1 : -1;
            packed.push_back(d);
        }
        if (int rc = s->d_ports.ensure(packed.size() * sizeof(PortDev))) return rc;
        if (int rc = s->d_phasors.ensure(packed.size() * 8 * sizeof(float))) return rc;
        EMFD_HIP(hipMemcpy(s->d_ports.ptr, packed.data(), packed.size() * sizeof(PortDev), hipMemcpyHostToDevice));
        EMFD_HIP(hipMemset(s->d_phasors.ptr, 0, packed.size() * 8 * sizeof(float)));
    }
    EMFD_HIP(hipStreamSynchronize(b->stream));
    if (applied) *applied = s->cfg;
    b->total_cells += static_cast<int64_t>(s->n);
    if (index) *index = static_cast<int>(b->scenes.size());
    b->scenes.push_back(std::move(s));
    return EMFD_OK;
}

int emfd_batch_finalize(emfd_batch *b, int64_t max_steps) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (b->finalized) return fail(EMFD_ERR_STATE, "batch already finalized");
    if (b->scenes.empty()) return fail(EMFD_ERR_STATE, "batch has no scenes");
    if (max_steps <= 0) return fail(EMFD_ERR_INVALID, "max_steps must be positive");
    EMFD_HIP(hipSetDevice(b->device));
    // step-row layout: every scene's signals, then a rotation quadruple per scene with ports
    int64_t off = 0;
    for (auto &s : b->scenes) {
        s->sig_offset = static_cast<int>(off);
        off += static_cast<int64_t>(s->sources.size());
    }
    for (auto &s : b->scenes) {
        if (s->ports.empty()) continue;
        s->rot_offset = static_cast<int>(off);
        off += 4;
        if (s->has_sweep && static_cast<size_t>(max_steps) > kSeriesFloats / (2 * s->ports.size()))
            return fail(EMFD_ERR_INVALID, "max_steps exceeds the port series capacity");
    }
    b->stride = std::max<int64_t>(off, 1);
    const size_t nfl = static_cast<size_t>(b->stride) * static_cast<size_t>(max_steps);
    std::vector<float> consts(nfl, 0.0f);
    for (auto &s : b->scenes) {
        const double dt = static_cast<double>(s->cfg.courant);
        for (int64_t k = 0; k < max_steps; ++k) {
            float *row = consts.data() + static_cast<size_t>(k) * b->stride;
            const double t = (static_cast<double>(k) + 0.5) * dt;
            for (size_t i = 0; i < s->sources.size(); ++i)
                row[s->sig_offset + static_cast<int64_t>(i)] = static_cast<float>(source_signal(s->sources[i], t));
            if (!s->ports.empty())
                port_rotation(s->wavelength, s->cfg.courant, static_cast<uint64_t>(k) + 1, row + s->rot_offset);
        }
    }
    if (int rc = b->d_consts.ensure(nfl * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpy(b->d_consts.ptr, consts.data(), nfl * sizeof(float), hipMemcpyHostToDevice));

    // series, descriptors, flat lists
    std::vector<BatchScene> desc;
    std::vector<int2> src_list, port_list;
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        SceneHost &s = *b->scenes[i];
        if (s.has_sweep && !s.ports.empty()) {
            const size_t bytes = static_cast<size_t>(max_steps) * s.ports.size() * 2 * sizeof(float);
            if (int rc = s.d_series.ensure(bytes)) return rc;
            EMFD_HIP(hipMemset(s.d_series.ptr, 0, bytes));
        }
        BatchScene d{};
        d.lat = s.lattice(b->pml_tabs[static_cast<size_t>(s.pml_table)].as<float>());
        d.f = s.field_ptrs();
        d.nq = static_cast<uint32_t>(s.n / 4);
        d.qx = static_cast<uint32_t>(s.cfg.size[0] / 4);
        d.mq = magic_of(d.qx);
        d.my = magic_of(static_cast<uint32_t>(s.cfg.size[1]));
        d.src_cells = s.d_cells.as<int>();
        d.src_lists = s.d_lists.as<int>();
        d.src_packed = s.d_packed.as<float>();
        d.sig_offset = s.sig_offset;
        d.n_ports = static_cast<int>(s.ports.size());
        d.ports = s.d_ports.as<PortDev>();
        d.phasors = s.d_phasors.as<float>();
        d.series = s.has_sweep && !s.ports.empty() ? s.d_series.as<float>() : nullptr;
        d.rot_offset = s.rot_offset;
        d.blend = s.blend;
        desc.push_back(d);
        if (!s.sources.empty())
            for (int c = 0; c < s.n_src_cells; ++c) src_list.push_back(make_int2(static_cast<int>(i), c));
        for (size_t p = 0; p < s.ports.size(); ++p) port_list.push_back(make_int2(static_cast<int>(i), static_cast<int>(p)));
    }
    if (int rc = b->d_scenes.ensure(desc.size() * sizeof(BatchScene))) return rc;
    EMFD_HIP(hipMemcpy(b->d_scenes.ptr, desc.data(), desc.size() * sizeof(BatchScene), hipMemcpyHostToDevice));
    b->n_src = static_cast<int>(src_list.size());
    b->n_port_items = static_cast<int>(port_list.size());
    if (int rc = b->d_src_list.ensure(std::max<size_t>(src_list.size(), 1) * sizeof(int2))) return rc;
    if (int rc = b->d_port_list.ensure(std::max<size_t>(port_list.size(), 1) * sizeof(int2))) return rc;
    if (!src_list.empty())
        EMFD_HIP(hipMemcpy(b->d_src_list.ptr, src_list.data(), src_list.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!port_list.empty())
        EMFD_HIP(hipMemcpy(b->d_port_list.ptr, port_list.data(), port_list.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (int rc = b->build_tiles()) return rc;
    b->max_steps = max_steps;
    b->finalized = true;
    return EMFD_OK;
}

int emfd_batch_info(emfd_batch *b, int *n_scenes, int64_t *total_cells, int64_t *n_tiles, size_t *device_bytes) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (n_scenes) *n_scenes = static_cast<int>(b->scenes.size());
    if (total_cells) *total_cells = b->total_cells;
    if (n_tiles) *n_tiles = b->n_tiles;
    if (device_bytes) {
        size_t t = 0;
        for (auto &s : b->scenes)
            for (const DevBuf *d : {&s->fields, &s->ids, &s->table, &s->prof, &s->d_packed, &s->d_cells, &s->d_lists,
                                    &s->d_ports, &s->d_phasors, &s->d_series})
                t += d->bytes;
        for (auto &p : b->pml_tabs) t += p.bytes;
        for (const DevBuf *d : {&b->d_scenes, &b->d_tiles, &b->d_src_list, &b->d_port_list, &b->d_consts, &b->d_step})
            t += d->bytes;
        *device_bytes = t;
    }
    return EMFD_OK;
}

int emfd_batch_set_launch(emfd_batch *b, const emfd_batch_launch *l) {
    if (!b || !l) return fail(EMFD_ERR_INVALID, "null argument");
    if (l->quads_per_thread != 1 && l->quads_per_thread != 2 && l->quads_per_thread != 4)
        return fail(EMFD_ERR_INVALID, "quads_per_thread must be 1, 2 or 4");
    if (l->xcd_slices != 1 && l->xcd_slices != 8) return fail(EMFD_ERR_INVALID, "xcd_slices must be 1 or 8");
    if (l->persistent_blocks < 0 || l->persistent_blocks > (1 << 20))
        return fail(EMFD_ERR_INVALID, "persistent_blocks out of range");
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    const bool retile = l->quads_per_thread != b->launch.quads_per_thread;
    b->launch = *l;
    b->drop_graph();
    if (retile && b->finalized) return b->build_tiles();
    return EMFD_OK;
}

int emfd_batch_get_launch(emfd_batch *b, emfd_batch_launch *l) {
    if (!b || !l) return fail(EMFD_ERR_INVALID, "null argument");
    *l = b->launch;
    return EMFD_OK;
}

int emfd_batch_autotune(emfd_batch *b, int iterations, emfd_batch_launch *chosen, double *ms_per_step, int cap,
                        int *n_tried) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (!b->finalized) return fail(EMFD_ERR_STATE, "batch not finalized");
    if (b->step_count != 0) return fail(EMFD_ERR_STATE, "autotune runs before the first step only");
    iterations = std::clamp(iterations, 1, 1000);
    EMFD_HIP(hipSetDevice(b->device));
    std::vector<emfd_batch_launch> cands;
    for (int q : {1, 2, 4})
        for (int pers : {0, 8 * b->cus})
            for (int sl : {1, 8}) cands.push_back({q, pers, sl});
    hipEvent_t e0, e1;
    EMFD_HIP(hipEventCreate(&e0));
    EMFD_HIP(hipEventCreate(&e1));
    double best = 1e300;
    emfd_batch_launch best_l = b->launch;
    int tried = 0;
    for (const emfd_batch_launch &l : cands) {
        if (int rc = emfd_batch_set_launch(b, &l)) return rc;
        const BatchLaunchArgs la = b->launch_args(l);
        EMFD_HIP(launch_batch_yee(la, l.persistent_blocks, b->stream));  // warm-up
        EMFD_HIP(hipEventRecord(e0, b->stream));
        for (int i = 0; i < iterations; ++i) EMFD_HIP(launch_batch_yee(la, l.persistent_blocks, b->stream));
        EMFD_HIP(hipEventRecord(e1, b->stream));
        EMFD_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        EMFD_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double per = static_cast<double>(ms) / iterations;
        if (ms_per_step && tried < cap) {
            ms_per_step[3 * tried + 0] = l.quads_per_thread + 0.01 * l.xcd_slices;
            ms_per_step[3 * tried + 1] = l.persistent_blocks;
            ms_per_step[3 * tried + 2] = per;
        }
        ++tried;
        if (per < best) {
            best = per;
            best_l = l;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (int rc = emfd_batch_set_launch(b, &best_l)) return rc;
    if (chosen) *chosen = best_l;
    if (n_tried) *n_tried = tried;
    return EMFD_OK;
}

int emfd_batch_set_graph(emfd_batch *b, int steps_per_graph) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (steps_per_graph < 0 || steps_per_graph > 4096) return fail(EMFD_ERR_INVALID, "steps_per_graph out of range");
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    b->drop_graph();
    b->graph_steps = steps_per_graph;
    return EMFD_OK;
}

int emfd_batch_step(emfd_batch *b, int count) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (!b->finalized) return fail(EMFD_ERR_STATE, "batch not finalized");
    if (count <= 0) return EMFD_OK;
    if (static_cast<int64_t>(b->step_count) + count > b->max_steps)
        return fail(EMFD_ERR_INVALID, "step count would exceed the batch's max_steps");
    EMFD_HIP(hipSetDevice(b->device));
    auto timed = [&](int launches, auto &&body) -> int {
        hipEvent_t a = nullptr, c = nullptr;
        if (b->timing) {
            if (b->ev_used + 2 > b->ev_pool.size()) {
                hipEvent_t x, y;
                EMFD_HIP(hipEventCreate(&x));
                EMFD_HIP(hipEventCreate(&y));
                b->ev_pool.push_back(x);
                b->ev_pool.push_back(y);
            }
            a = b->ev_pool[b->ev_used];
            c = b->ev_pool[b->ev_used + 1];
            b->ev_used += 2;
            b->ev_launches.push_back(launches);
            EMFD_HIP(hipEventRecord(a, b->stream));
        }
        if (int rc = body()) return rc;
        if (c) EMFD_HIP(hipEventRecord(c, b->stream));
        if (b->timing && b->ev_used >= 4096) return b->collect_timing();
        return EMFD_OK;
    };
    int done = 0;
    if (b->graph_steps > 0) {
        while (count - done >= b->graph_steps) {
            if (!b->graph_exec) {
                EMFD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                const int rc = b->enqueue_steps(b->graph_steps);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(b->stream, &g);
                if (rc) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                EMFD_HIP(ec);
                b->graph = g;
                EMFD_HIP(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
            }
            if (int rc = timed(b->graph_steps, [&]() -> int {
                    EMFD_HIP(hipGraphLaunch(b->graph_exec, b->stream));
                    return EMFD_OK;
                }))
                return rc;
            done += b->graph_steps;
        }
    }
    if (count > done) {
        // direct launches; with timing on, each step's two Yee passes are bracketed
        const BatchLaunchArgs la = b->launch_args(b->launch);
        const float *consts = b->d_consts.as<float>();
        int64_t *ds = b->d_step.as<int64_t>();
        const int rest = count - done;
        for (int k = 0; k < rest; ++k) {
            if (int rc = timed(1, [&]() -> int {
                    EMFD_HIP(launch_batch_yee(la, b->launch.persistent_blocks, b->stream));
                    return EMFD_OK;
                }))
                return rc;
            EMFD_HIP(launch_batch_sources(b->d_scenes.as<BatchScene>(), b->d_src_list.as<int2>(), b->n_src, consts,
                                          b->stride, ds, k, b->stream));
            EMFD_HIP(launch_batch_ports(b->d_scenes.as<BatchScene>(), b->d_port_list.as<int2>(), b->n_port_items,
                                        consts, b->stride, ds, k, b->stream));
        }
        EMFD_HIP(launch_batch_advance(ds, rest, b->stream));
    }
    b->step_count += static_cast<uint64_t>(count);
    for (auto &s : b->scenes) s->evaluated = false;
    return EMFD_OK;
}

int emfd_batch_synchronize(emfd_batch *b) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    return EMFD_OK;
}

int emfd_batch_state(emfd_batch *b, uint64_t *step_count) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (step_count) *step_count = b->step_count;
    return EMFD_OK;
}

int emfd_batch_reset(emfd_batch *b) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    for (auto &s : b->scenes) {
        EMFD_HIP(hipMemsetAsync(s->fields.ptr, 0, s->n * 6 * sizeof(float), b->stream));
        if (s->d_phasors.ptr) EMFD_HIP(hipMemsetAsync(s->d_phasors.ptr, 0, s->d_phasors.bytes, b->stream));
        if (s->d_series.ptr) EMFD_HIP(hipMemsetAsync(s->d_series.ptr, 0, s->d_series.bytes, b->stream));
        s->evaluated = false;
    }
    EMFD_HIP(hipMemsetAsync(b->d_step.ptr, 0, sizeof(int64_t), b->stream));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    b->step_count = 0;
    return EMFD_OK;
}

static int scene_of(emfd_batch *b, int scene, SceneHost **out) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (scene < 0 || scene >= static_cast<int>(b->scenes.size())) return fail(EMFD_ERR_INVALID, "no such scene");
    *out = b->scenes[static_cast<size_t>(scene)].get();
    return EMFD_OK;
}

int emfd_batch_download_fields(emfd_batch *b, int scene, float *E, float *B) {
    SceneHost *s = nullptr;
    if (int rc = scene_of(b, scene, &s)) return rc;
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    if (E) EMFD_HIP(hipMemcpy(E, s->fields.ptr, s->n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    if (B) EMFD_HIP(hipMemcpy(B, s->fields.as<float>() + 3 * s->n, s->n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return EMFD_OK;
}

int emfd_batch_upload_fields(emfd_batch *b, int scene, const float *E, const float *B) {
    SceneHost *s = nullptr;
    if (int rc = scene_of(b, scene, &s)) return rc;
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    if (E) EMFD_HIP(hipMemcpy(s->fields.ptr, E, s->n * 3 * sizeof(float), hipMemcpyHostToDevice));
    if (B) EMFD_HIP(hipMemcpy(s->fields.as<float>() + 3 * s->n, B, s->n * 3 * sizeof(float), hipMemcpyHostToDevice));
    return EMFD_OK;
}

int emfd_batch_read_phasors(emfd_batch *b, int scene, float *out, int n) {
    SceneHost *s = nullptr;
    if (int rc = scene_of(b, scene, &s)) return rc;
    if (!out) return fail(EMFD_ERR_INVALID, "null output");
    const size_t need = s->ports.size() * 8;
    if (n < 0 || static_cast<size_t>(n) < need) return fail(EMFD_ERR_INVALID, "output too small");
    if (need == 0) return EMFD_OK;
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    EMFD_HIP(hipMemcpy(out, s->d_phasors.ptr, need * sizeof(float), hipMemcpyDeviceToHost));
    return EMFD_OK;
}

int emfd_batch_read_series(emfd_batch *b, int scene, float *out, int64_t n) {
    SceneHost *s = nullptr;
    if (int rc = scene_of(b, scene, &s)) return rc;
    if (!out) return fail(EMFD_ERR_INVALID, "null output");
    if (!s->d_series.ptr) return fail(EMFD_ERR_STATE, "scene records no series (no sweep)");
    const size_t need = static_cast<size_t>(b->step_count) * s->ports.size() * 2;
    if (n < 0 || static_cast<size_t>(n) < need) return fail(EMFD_ERR_INVALID, "output too small");
    if (need == 0) return EMFD_OK;
    EMFD_HIP(hipSetDevice(b->device));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    EMFD_HIP(hipMemcpy(out, s->d_series.ptr, need * sizeof(float), hipMemcpyDeviceToHost));
    return EMFD_OK;
}

int emfd_batch_evaluate_sweeps(emfd_batch *b) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    if (!b->finalized) return fail(EMFD_ERR_STATE, "batch not finalized");
    EMFD_HIP(hipSetDevice(b->device));
    const int64_t n = static_cast<int64_t>(b->step_count);
    std::vector<int> which;
    std::vector<SweepBand> bands;
    size_t n_steps = 0, n_sums = 0;
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        SceneHost &s = *b->scenes[i];
        s.evaluated = false;
        if (!s.has_sweep || s.ports.empty() || n < 16) continue;
        which.push_back(static_cast<int>(i));
        bands.push_back(sweep_band(s.sweep));
        n_steps += static_cast<size_t>(bands.back().points);
        n_sums += static_cast<size_t>(bands.back().points) * s.ports.size() * 2;
    }
    if (which.empty()) return EMFD_OK;
    std::vector<double2> steps(n_steps);
    std::vector<BatchSweepJob> jobs(which.size());
    if (int rc = b->d_sweep_steps.ensure(n_steps * sizeof(double2))) return rc;
    if (int rc = b->d_sums.ensure(n_sums * sizeof(double2))) return rc;
    if (int rc = b->d_residual.ensure(which.size() * sizeof(double))) return rc;
    if (int rc = b->d_jobs.ensure(which.size() * sizeof(BatchSweepJob))) return rc;
    size_t so = 0, mo = 0;
    std::vector<size_t> step_off(which.size()), sum_off(which.size());
    for (size_t j = 0; j < which.size(); ++j) {
        SceneHost &s = *b->scenes[static_cast<size_t>(which[j])];
        const double dt = static_cast<double>(s.cfg.courant);
        for (int f = 0; f < bands[j].points; ++f) {
            const std::complex<double> st = std::polar(1.0, -(kTwoPi * bands[j].nu(f)) * dt);
            steps[so + static_cast<size_t>(f)] = make_double2(st.real(), st.imag());
        }
        BatchSweepJob jb{};
        jb.series = s.d_series.as<float>();
        jb.n_ports = static_cast<int>(s.ports.size());
        jb.points = bands[j].points;
        jb.n = n;
        jb.step = b->d_sweep_steps.as<double2>() + so;
        jb.sums = b->d_sums.as<double2>() + mo;
        jb.residual = b->d_residual.as<double>() + j;
        jobs[j] = jb;
        step_off[j] = so;
        sum_off[j] = mo;
        so += static_cast<size_t>(bands[j].points);
        mo += static_cast<size_t>(bands[j].points) * s.ports.size() * 2;
    }
    EMFD_HIP(hipMemcpyAsync(b->d_sweep_steps.ptr, steps.data(), n_steps * sizeof(double2), hipMemcpyHostToDevice,
                            b->stream));
    EMFD_HIP(hipMemcpyAsync(b->d_jobs.ptr, jobs.data(), jobs.size() * sizeof(BatchSweepJob), hipMemcpyHostToDevice,
                            b->stream));
    EMFD_HIP(launch_batch_sweep(b->d_jobs.as<BatchSweepJob>(), static_cast<int>(jobs.size()), b->stream));
    std::vector<double2> sums(n_sums);
    std::vector<double> resid(which.size());
    EMFD_HIP(hipMemcpyAsync(sums.data(), b->d_sums.ptr, n_sums * sizeof(double2), hipMemcpyDeviceToHost, b->stream));
    EMFD_HIP(hipMemcpyAsync(resid.data(), b->d_residual.ptr, which.size() * sizeof(double), hipMemcpyDeviceToHost,
                            b->stream));
    EMFD_HIP(hipStreamSynchronize(b->stream));
    for (size_t j = 0; j < which.size(); ++j) {
        SceneHost &s = *b->scenes[static_cast<size_t>(which[j])];
        const SweepBand &band = bands[j];
        const double dt = static_cast<double>(s.cfg.courant);
        const size_t P = s.ports.size(), npts = static_cast<size_t>(band.points);
        s.points.assign(P * npts, emfd_sweep_point{});
        s.summary.assign(P, emfd_port_sweep{});
        std::vector<std::complex<double>> sv(npts), si(npts);
        for (size_t p = 0; p < P; ++p) {
            for (size_t f = 0; f < npts; ++f) {
                const double2 v = sums[sum_off[j] + 2 * (p * npts + f)], c = sums[sum_off[j] + 2 * (p * npts + f) + 1];
                const double w = kTwoPi * band.nu(static_cast<int>(f));
                sv[f] = std::complex<double>(v.x, v.y);
                si[f] = std::complex<double>(c.x, c.y) * std::polar(1.0, w * 0.5 * dt);
            }
            sweep_finish(band, sv.data(), si.data(), static_cast<double>(s.ports[p].impedance_ohm),
                         s.points.data() + p * npts, s.summary.data() + p);
        }
        s.residual = resid[j];
        s.evaluated = true;
    }
    return EMFD_OK;
}

int emfd_batch_sweep_result(emfd_batch *b, int scene, emfd_sweep_point *points, int cap_points,
                            emfd_port_sweep *summary, double *residual) {
    SceneHost *s = nullptr;
    if (int rc = scene_of(b, scene, &s)) return rc;
    if (!s->evaluated) return fail(EMFD_ERR_STATE, "no evaluated sweep for this scene");
    if (points) {
        if (cap_points < 0 || static_cast<size_t>(cap_points) < s->points.size())
            return fail(EMFD_ERR_INVALID, "points buffer too small");
        std::copy(s->points.begin(), s->points.end(), points);
    }
    if (summary) std::copy(s->summary.begin(), s->summary.end(), summary);
    if (residual) *residual = s->residual;
    return EMFD_OK;
}

int emfd_batch_kernel_timing(emfd_batch *b, int enable, double *total_ms, int64_t *launches) {
    if (!b) return fail(EMFD_ERR_INVALID, "null batch");
    EMFD_HIP(hipSetDevice(b->device));
    if (int rc = b->collect_timing()) return rc;
    if (total_ms) *total_ms = b->timed_ms;
    if (launches) *launches = b->timed_launches;
    if (enable >= 0) {
        b->timing = enable != 0;
        b->timed_ms = 0.0;
        b->timed_launches = 0;
    }
    return EMFD_OK;
}

}  // extern "C"
