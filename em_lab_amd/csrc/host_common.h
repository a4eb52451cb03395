// Host-side pieces of SimulationGrid shared by the single-grid context (grid.cpp) and the batched
// stepper (batch.cpp): configure clamps, absorber profiles, material classification (edge
// averages, sub-cell overrides, binary16 keys) and source packing. Not part of the public ABI.
#pragma once

#include "../../include/emfdtd.h"

#include <array>
#include <complex>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace emfd {

uint16_t f32_to_f16(float f);
double source_signal(const emfd_source_desc &s, double t);

// SimulationGrid::configure clamps (src/SimulationGrid.cpp:150-156)
emfd_grid_config clamp_config(const emfd_grid_config &requested);

// pmlProfile(depth) per axis at i and i + 1/2 (shaders/common.glsl:23-36): 6 x stride floats
std::vector<float> absorber_profiles(const emfd_grid_config &c, int *stride);

// Coefficient classes: 12 binary16 values per texel (edge sigma x y z clamped to 6e4, edge eps_r
// x y z, thin-wire factors of Bx, By, Bz terms 0 and 1). Class 0 is vacuum.
using ClassKey = std::array<uint16_t, 12>;
struct ClassKeyHash {
    size_t operator()(const ClassKey &k) const noexcept {
        uint64_t h = 1469598103934665603ull;
        for (uint16_t v : k) h = (h ^ v) * 1099511628211ull;
        return static_cast<size_t>(h);
    }
};
struct ClassMap {
    std::unordered_map<ClassKey, uint32_t, ClassKeyHash> map;
    std::vector<ClassKey> keys;
    ClassMap();
    static ClassKey vacuum();
};

// uploadMaterials' edge averaging and overrides (src/SimulationGrid.cpp:224-291), quantised and
// classified; new classes are appended to `cm`. ids: one per texel.
void classify_cells(const emfd_grid_config &c, const float *sigma, const float *eps,
                    const emfd_edge_override *edges, int64_t n_edges, const emfd_wire_coefficient *wires,
                    int64_t n_wires, ClassMap &cm, uint32_t *ids);

// Whether some texel of class != 0 lies in the absorber (the texels in_pml() marks): the E pass
// then needs its general coefficient path.
bool absorber_has_materials(const emfd_grid_config &c, const uint32_t *ids);

// setSources (src/SimulationGrid.cpp:297-357): packed 16 floats per kept source, per-texel lists
struct PackedSources {
    std::vector<float> packed;
    std::vector<emfd_source_desc> kept;
    std::vector<int> cells;  // 4 ints per texel (x, y, z, first) + sentinel
    std::vector<int> lists;
    int n_cells = 0;
};
PackedSources pack_sources(const emfd_grid_config &c, const emfd_source_desc *src, int n);

// PortMonitor sweep (src/Ports.cpp:187-218): nu evenly spaced over [1/lambda_max, 1/lambda_min]
struct SweepBand {
    int points = 2;
    double nu_lo = 0.0, nu_hi = 0.0;
    double nu(int f) const { return nu_lo + (nu_hi - nu_lo) * f / (points - 1); }
};
SweepBand sweep_band(const emfd_sweep_setup &sweep);
// Z, S11, VSWR per point and the best/band summary of one port from its DFT sums (sumI already
// carries the e^{+i w dt / 2} half-step factor). points: band.points entries, or null.
void sweep_finish(const SweepBand &band, const std::complex<double> *sumV, const std::complex<double> *sumI,
                  double impedance, emfd_sweep_point *points, emfd_port_sweep *summary);
// Rotation factors (cos w tE, sin w tE, cos w tH, sin w tH) at step_after, FP64 cast to float,
// and the per-stage EMA weight of the port and far-field accumulators.
void port_rotation(float wavelength, float dt, uint64_t step_after, float rot[4]);
float port_blend(float averaging, float wavelength, float dt);

}  // namespace emfd
