// The binding a maintainer of EM_Lab would add in place of the OpenGL bodies of SimulationGrid,
// PortMonitor and FarFieldMonitor (src/SimulationGrid.hpp:155-216, src/Ports.hpp:84-140,
// src/FarField.hpp:75-133), compiled against include/emfdtd.h and libemfdtd.so. The reference's
// own types are restated here in the shape the binding reads (glm vectors as plain structs), so
// that the binding compiles as it would inside EM_Lab.
//
// As a program it runs one scene bundle through the binding and writes what the solver produced:
//
//   simulation_grid <bundle dir> <out dir>
//
// bundle: grid.bin (emfd_grid_config), sigma.f32 / eps.f32 (per cell), edges.bin, wires.bin,
// sources.bin, ports.bin (arrays of the ABI structs), sweep.bin (emfd_sweep_setup, optional),
// farfield.bin (emfd_far_field_setup, optional), run.txt ("steps averaging_periods cell_size"),
// camera.f32 (optional: a frame of the volume view).
// out: E.f32 / B.f32 (3 x n component planes), port_phasors.f32, port_series.f32, sweep.f64
// (wavelength, R, X, S11 dB, VSWR per port and point), readings.f64, ntff_phasors.f32,
// pattern.txt (valid, radiated power, max directivity).
#include "emfdtd.h"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

// ---- the reference's types, as the binding sees them ----
struct ivec3 { int x, y, z; };
struct vec3 { float x, y, z; };

struct GridConfig {  // src/SimulationGrid.hpp:25-31
    ivec3 size;
    float courant;
    int pmlThickness;
    float pmlSigmaMax;
    float pmlGrading;
};
struct EdgeOverride { ivec3 texel; int axis; float sigma; };                        // :76-81
struct WireCoefficient { ivec3 texel; int component; int term; float factor; };      // :86-94
struct SubcellModel { std::vector<EdgeOverride> edges; std::vector<WireCoefficient> wireTerms; };
class MaterialVolume {
public:
    std::vector<float> s, e;
    const std::vector<float> &sigma() const { return s; }
    const std::vector<float> &epsilonR() const { return e; }
};
enum class Waveform : int { CW = 0, Pulse = 1 };
struct SourceDesc {  // src/SimulationGrid.hpp:113-150
    vec3 center, halfExtent, softness, polarization;
    float amplitude, wavelength, phase;
    Waveform waveform;
    float rampPeriods, repeatPeriods, delayPeriods;
    bool enabled;
};
struct PortSetup {  // src/Ports.hpp:30-40
    ivec3 texel;
    int axis;
    float sign;
    int edges;
    float wavelength;
    bool voltageFeed;
    float impedanceOhm;
    float drive;
};
struct SweepSetup { float wavelengthMin, wavelengthMax; int points; };  // src/Ports.hpp:42-47
struct FarFieldSetup {                                                   // src/FarField.hpp:39-47
    ivec3 boxMin, boxMax;
    float wavelength, averagingPeriods;
    vec3 mapAxis;
};

static void check(int rc) {
    if (rc != EMFD_OK) throw std::runtime_error(emfd_last_error());
}

// ---- SimulationGrid ----
class SimulationGrid {
public:
    SimulationGrid() { check(emfd_create(0, &m_ctx)); }
    ~SimulationGrid() { emfd_destroy(m_ctx); }
    SimulationGrid(const SimulationGrid &) = delete;
    SimulationGrid &operator=(const SimulationGrid &) = delete;

    void configure(const GridConfig &c) {
        emfd_grid_config in{{c.size.x, c.size.y, c.size.z}, c.courant, c.pmlThickness, c.pmlSigmaMax, c.pmlGrading},
            out{};
        check(emfd_grid_configure(m_ctx, &in, &out));
        m_config = {{out.size[0], out.size[1], out.size[2]}, out.courant, out.pml_thickness, out.pml_sigma_max,
                    out.pml_grading};
    }

    void uploadMaterials(const MaterialVolume &m, const SubcellModel &s) {
        static_assert(sizeof(EdgeOverride) == sizeof(emfd_edge_override), "ivec3 + int + float");
        static_assert(sizeof(WireCoefficient) == sizeof(emfd_wire_coefficient), "ivec3 + int + int + float");
        check(emfd_grid_upload_materials(m_ctx, m.sigma().data(), m.epsilonR().data(),
                                         static_cast<int64_t>(m.sigma().size()),
                                         reinterpret_cast<const emfd_edge_override *>(s.edges.data()),
                                         static_cast<int64_t>(s.edges.size()),
                                         reinterpret_cast<const emfd_wire_coefficient *>(s.wireTerms.data()),
                                         static_cast<int64_t>(s.wireTerms.size())));
    }

    void setSources(const std::vector<SourceDesc> &src) {
        std::vector<emfd_source_desc> v;
        for (const SourceDesc &s : src)
            v.push_back({{s.center.x, s.center.y, s.center.z},
                         {s.halfExtent.x, s.halfExtent.y, s.halfExtent.z},
                         {s.softness.x, s.softness.y, s.softness.z},
                         {s.polarization.x, s.polarization.y, s.polarization.z},
                         s.amplitude, s.wavelength, s.phase, static_cast<int>(s.waveform), s.rampPeriods,
                         s.repeatPeriods, s.delayPeriods, s.enabled ? 1 : 0});
        int kept = 0;
        check(emfd_grid_set_sources(m_ctx, v.data(), static_cast<int>(v.size()), &kept));
    }

    void step(int count, const std::function<void()> &afterEachStep = {}) {
        if (!afterEachStep) {
            check(emfd_grid_step(m_ctx, count));
            return;
        }
        for (int k = 0; k < count; ++k) {
            check(emfd_grid_step(m_ctx, 1));
            afterEachStep();
        }
    }

    double dt() const {
        double t = 0.0;
        uint64_t n = 0;
        float dt = 0.0f;
        check(emfd_grid_state(m_ctx, &t, &n, &dt));
        return dt;
    }
    const GridConfig &config() const { return m_config; }
    emfd_ctx *context() const { return m_ctx; }

private:
    emfd_ctx *m_ctx = nullptr;
    GridConfig m_config{};
};

// ---- PortMonitor ----
class PortMonitor {
public:
    explicit PortMonitor(SimulationGrid &g) : m_grid(g) {}
    void configure(const std::vector<PortSetup> &ports, float averagingPeriods, const SweepSetup *sweep,
                   float cellSize) {
        std::vector<emfd_port_setup> v;
        for (const PortSetup &p : ports)
            v.push_back({{p.texel.x, p.texel.y, p.texel.z}, p.axis, p.sign, p.edges, p.wavelength,
                         p.voltageFeed ? 1 : 0, p.impedanceOhm, p.drive});
        emfd_sweep_setup sw{};
        if (sweep) sw = {sweep->wavelengthMin, sweep->wavelengthMax, sweep->points};
        check(emfd_ports_configure(m_grid.context(), v.data(), static_cast<int>(v.size()), averagingPeriods,
                                   sweep ? &sw : nullptr, cellSize, 0));
        m_ports = ports.size();
        m_points = sweep ? sweep->points : 0;
    }
    // accumulate() is a no-op: emfd_grid_step runs the port step after every step
    void accumulate() {}
    std::vector<emfd_port_reading> evaluate() const {
        std::vector<emfd_port_reading> r(m_ports);
        check(emfd_ports_evaluate(m_grid.context(), r.data(), static_cast<int>(m_ports)));
        return r;
    }
    std::vector<emfd_sweep_point> evaluateSweep(double dt) const {
        std::vector<emfd_sweep_point> pts(m_ports * static_cast<size_t>(m_points));
        std::vector<emfd_port_sweep> summary(m_ports);
        double residual = 0.0;
        int64_t samples = 0;
        int full = 0;
        check(emfd_ports_evaluate_sweep(m_grid.context(), dt, pts.data(), summary.data(), &residual, &samples,
                                        &full));
        return pts;
    }
    std::vector<float> phasors() const {
        std::vector<float> out(8 * m_ports);
        check(emfd_ports_read_phasors(m_grid.context(), out.data(), static_cast<int>(m_ports)));
        return out;
    }
    std::vector<float> series() const {
        const int64_t n = emfd_ports_recorded_steps(m_grid.context());
        std::vector<float> out(static_cast<size_t>(n) * m_ports * 2);
        check(emfd_ports_read_series(m_grid.context(), out.data(), static_cast<int64_t>(out.size())));
        return out;
    }
    size_t size() const { return m_ports; }

private:
    SimulationGrid &m_grid;
    size_t m_ports = 0;
    int m_points = 0;
};

// ---- FarFieldMonitor ----
class FarFieldMonitor {
public:
    explicit FarFieldMonitor(SimulationGrid &g) : m_grid(g) {}
    void configure(const FarFieldSetup &s) {
        emfd_far_field_setup in{{s.boxMin.x, s.boxMin.y, s.boxMin.z},
                                {s.boxMax.x, s.boxMax.y, s.boxMax.z},
                                s.wavelength,
                                s.averagingPeriods,
                                {s.mapAxis.x, s.mapAxis.y, s.mapAxis.z}},
            applied{};
        check(emfd_ntff_configure(m_grid.context(), &in, &applied, &m_points));
    }
    void accumulate() {}
    std::vector<float> phasors() const {
        std::vector<float> out(static_cast<size_t>(m_points) * 32);
        check(emfd_ntff_read_phasors(m_grid.context(), out.data(), m_points));
        return out;
    }
    // FarFieldMonitor::evaluate + finishEvaluation: valid, radiated power, peak directivity
    void evaluate(int *valid, double *power, float *maxDbi) const {
        float dir[3];
        check(emfd_ntff_evaluate(m_grid.context(), valid, power, maxDbi, dir, nullptr, nullptr));
    }

private:
    SimulationGrid &m_grid;
    int m_points = 0;
};

// ---- VolumeRenderer (src/VolumeRenderer.cpp:70-121): the volume pass as one call ----
struct RenderSettings {  // src/VolumeRenderer.hpp:21-39 (the fields the volume pass reads)
    int mode = 0, signedAxis = 2;
    float gainE = 1.0f, gainB = 1.0f, density = 1.0f, emission = 1.0f, stepVoxels = 0.75f;
    float materialOpacity = 0.45f, exposure = 1.25f;
    float clipMin[3] = {0.0f, 0.0f, 0.0f}, clipMax[3] = {1.0f, 1.0f, 1.0f};
};
class VolumeRenderer {
public:
    // invViewProj: inverse(camera.viewProjection()), column-major; eye: camera.position()
    std::vector<uint8_t> render(const SimulationGrid &grid, const float *invViewProj, const float *eye,
                                const RenderSettings &rs, int width, int height) const {
        emfd_render_settings s;
        emfd_render_default_settings(&s);
        s.mode = rs.mode;
        s.signed_axis = rs.signedAxis;
        s.gain_e = rs.gainE;
        s.gain_b = rs.gainB;
        s.density = rs.density;
        s.emission = rs.emission;
        s.step_voxels = rs.stepVoxels;
        s.material_opacity = rs.materialOpacity;
        s.exposure = rs.exposure;
        for (int a = 0; a < 3; ++a) {
            s.clip_min[a] = rs.clipMin[a];
            s.clip_max[a] = rs.clipMax[a];
        }
        std::vector<uint8_t> rgba(static_cast<size_t>(width) * height * 4);
        check(emfd_render_volume(grid.context(), &s, invViewProj, eye, width, height, EMFD_IMAGE_RGBA8, rgba.data()));
        return rgba;
    }
};

// ---- the program ----
template <class T>
static std::vector<T> read_array(const std::string &path, bool required) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) {
        if (required) throw std::runtime_error("cannot open " + path);
        return {};
    }
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    if (bytes % static_cast<std::streamsize>(sizeof(T)) != 0) throw std::runtime_error("bad size: " + path);
    std::vector<T> v(static_cast<size_t>(bytes) / sizeof(T));
    if (bytes > 0) f.read(reinterpret_cast<char *>(v.data()), bytes);
    return v;
}

template <class T>
static void write_array(const std::string &path, const std::vector<T> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
    if (!f) throw std::runtime_error("cannot write " + path);
}

int main(int argc, char **argv) {
    if (argc != 3) {
        std::cerr << "usage: simulation_grid <bundle dir> <out dir>\n";
        return 2;
    }
    const std::string in = argv[1], out = argv[2];
    try {
        const auto gc = read_array<emfd_grid_config>(in + "/grid.bin", true);
        if (gc.size() != 1) throw std::runtime_error("grid.bin holds one emfd_grid_config");
        GridConfig cfg{{gc[0].size[0], gc[0].size[1], gc[0].size[2]}, gc[0].courant, gc[0].pml_thickness,
                       gc[0].pml_sigma_max, gc[0].pml_grading};
        MaterialVolume mat;
        mat.s = read_array<float>(in + "/sigma.f32", true);
        mat.e = read_array<float>(in + "/eps.f32", true);
        SubcellModel sub;
        for (const auto &e : read_array<emfd_edge_override>(in + "/edges.bin", false))
            sub.edges.push_back({{e.texel[0], e.texel[1], e.texel[2]}, e.axis, e.sigma});
        for (const auto &w : read_array<emfd_wire_coefficient>(in + "/wires.bin", false))
            sub.wireTerms.push_back({{w.texel[0], w.texel[1], w.texel[2]}, w.component, w.term, w.factor});
        std::vector<SourceDesc> sources;
        for (const auto &s : read_array<emfd_source_desc>(in + "/sources.bin", false))
            sources.push_back({{s.center[0], s.center[1], s.center[2]},
                               {s.half_extent[0], s.half_extent[1], s.half_extent[2]},
                               {s.softness[0], s.softness[1], s.softness[2]},
                               {s.polarization[0], s.polarization[1], s.polarization[2]},
                               s.amplitude, s.wavelength, s.phase, static_cast<Waveform>(s.waveform),
                               s.ramp_periods, s.repeat_periods, s.delay_periods, s.enabled != 0});
        std::vector<PortSetup> ports;
        for (const auto &p : read_array<emfd_port_setup>(in + "/ports.bin", false))
            ports.push_back({{p.texel[0], p.texel[1], p.texel[2]}, p.axis, p.sign, p.edges, p.wavelength,
                             p.voltage_feed != 0, p.impedance_ohm, p.drive});
        const auto sw = read_array<emfd_sweep_setup>(in + "/sweep.bin", false);
        const auto ff = read_array<emfd_far_field_setup>(in + "/farfield.bin", false);
        int steps = 0;
        float averaging = 4.0f, cell = 1e-3f;
        {
            std::ifstream r(in + "/run.txt");
            if (!(r >> steps >> averaging >> cell)) throw std::runtime_error("run.txt: steps averaging cell_size");
        }

        SimulationGrid grid;
        grid.configure(cfg);
        grid.uploadMaterials(mat, sub);
        grid.setSources(sources);
        PortMonitor pm(grid);
        SweepSetup sweep{};
        if (!ports.empty()) {
            if (!sw.empty()) sweep = {sw[0].wavelength_min, sw[0].wavelength_max, sw[0].points};
            pm.configure(ports, averaging, sw.empty() ? nullptr : &sweep, cell);
        }
        FarFieldMonitor far(grid);
        if (!ff.empty())
            far.configure({{ff[0].box_min[0], ff[0].box_min[1], ff[0].box_min[2]},
                           {ff[0].box_max[0], ff[0].box_max[1], ff[0].box_max[2]},
                           ff[0].wavelength,
                           ff[0].averaging_periods,
                           {ff[0].map_axis[0], ff[0].map_axis[1], ff[0].map_axis[2]}});
        // the --frames loop's shape: steps in two calls, monitors attached to the context
        grid.step(steps / 2);
        grid.step(steps - steps / 2);

        const GridConfig &c = grid.config();
        const size_t n = static_cast<size_t>(c.size.x) * c.size.y * c.size.z;
        std::vector<float> E(3 * n), B(3 * n);
        check(emfd_grid_download_fields(grid.context(), E.data(), B.data(), EMFD_LAYOUT_SOA));
        write_array(out + "/E.f32", E);
        write_array(out + "/B.f32", B);
        if (pm.size() > 0) {
            write_array(out + "/port_phasors.f32", pm.phasors());
            write_array(out + "/port_series.f32", pm.series());
            std::vector<double> rd;
            for (const emfd_port_reading &r : pm.evaluate())
                for (double v : {static_cast<double>(r.valid), r.z_re, r.z_im, r.gamma_re, r.gamma_im, r.vswr,
                                 r.return_loss_db, r.accepted_power, r.available_power})
                    rd.push_back(v);
            write_array(out + "/readings.f64", rd);
            if (!sw.empty()) {
                std::vector<double> sp;
                for (const emfd_sweep_point &p : pm.evaluateSweep(grid.dt()))
                    for (double v : {p.wavelength, p.z_re, p.z_im, p.s11_db, p.vswr}) sp.push_back(v);
                write_array(out + "/sweep.f64", sp);
            }
        }
        if (!ff.empty()) {
            write_array(out + "/ntff_phasors.f32", far.phasors());
            int valid = 0;
            double power = 0.0;
            float dbi = 0.0f;
            far.evaluate(&valid, &power, &dbi);
            std::ofstream p(out + "/pattern.txt");
            char line[128];
            std::snprintf(line, sizeof line, "%d %.17g %.9g\n", valid, power, static_cast<double>(dbi));
            p << line;
        }
        // camera.f32 (optional): 16 floats of the inverse view-projection, 3 of the eye, then
        // width, height, mode as floats; the frame goes to frame.rgba
        const auto cam = read_array<float>(in + "/camera.f32", false);
        if (cam.size() == 22) {
            RenderSettings rs;
            rs.mode = static_cast<int>(cam[21]);
            write_array(out + "/frame.rgba", VolumeRenderer().render(grid, cam.data(), cam.data() + 16, rs,
                                                                      static_cast<int>(cam[19]), static_cast<int>(cam[20])));
        }
        std::cout << "simulation_grid: " << emfd_version() << " | " << c.size.x << "x" << c.size.y << "x" << c.size.z
                  << ", " << steps << " steps, " << pm.size() << " ports" << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "simulation_grid: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
