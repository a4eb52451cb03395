// libemfdtd.so host side of the monitors: PortMonitor (src/Ports.cpp), FarFieldMonitor
// (src/FarField.cpp) and FieldAnalyzer (src/FieldAnalyzer.cpp) entry points. Per-step work runs
// on the device inside emfd_grid_step; evaluation math that the reference does on the CPU in
// double precision stays on the host here, with the same formulas.
#include "context.h"
#include "host_common.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>

namespace emfd {
namespace {
constexpr double kEta0 = 376.730313668;  // src/Ports.cpp:153
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr int kMaxPorts = 256;                        // src/Ports.cpp:155
constexpr size_t kSeriesFloats = size_t(1) << 23;     // src/Ports.cpp:156
constexpr int kCutSamples = 180;                      // src/FarField.cpp:148
constexpr int kMapSize = 81;                          // src/FarField.cpp:149

std::complex<double> reflection(const std::complex<double> &z, double z0) { return (z - z0) / (z + z0); }
double vswr_of(double g) {
    g = std::min(g, 0.999999);
    return (1.0 + g) / (1.0 - g);
}

// rotation factors (cos w tE, sin w tE, cos w tH, sin w tH) for the step that leaves the grid
// at step_after, and the per-stage EMA weight (src/Ports.cpp:253-262, src/FarField.cpp:331-344)
void rotation(float wavelength, float dt, uint64_t step_after, float rot[4]) {
    const double w = kTwoPi / static_cast<double>(wavelength);
    const double tE = static_cast<double>(step_after) * static_cast<double>(dt);
    const double tH = tE - 0.5 * static_cast<double>(dt);
    rot[0] = static_cast<float>(std::cos(w * tE));
    rot[1] = static_cast<float>(std::sin(w * tE));
    rot[2] = static_cast<float>(std::cos(w * tH));
    rot[3] = static_cast<float>(std::sin(w * tH));
}
float blend_of(float averaging, float wavelength, float dt) {
    const double tau = averaging * wavelength;  // float product, as the reference
    return static_cast<float>(1.0 - std::exp(-static_cast<double>(dt) / (0.5 * tau)));
}

}  // namespace

SweepBand sweep_band(const emfd_sweep_setup &sw) {
    SweepBand b;
    b.points = std::max(sw.points, 2);
    b.nu_lo = 1.0 / std::max(sw.wavelength_max, sw.wavelength_min);
    b.nu_hi = 1.0 / std::min(sw.wavelength_max, sw.wavelength_min);
    return b;
}

void sweep_finish(const SweepBand &band, const std::complex<double> *sumV, const std::complex<double> *sumI,
                  double impedance, emfd_sweep_point *points, emfd_port_sweep *summary) {
    // src/Ports.cpp:234-278
    const double z0 = std::max(impedance, 1e-6);
    const int npts = band.points;
    std::vector<emfd_sweep_point> pts;
    for (int f = 0; f < npts; ++f) {
        emfd_sweep_point pt{};
        pt.wavelength = 1.0 / band.nu(f);
        const std::complex<double> &sv = sumV[f], &si = sumI[f];
        if (std::abs(si) > 1e-30) {
            const std::complex<double> z = -sv / si * kEta0;
            const double g = std::abs(reflection(z, z0));
            pt.z_re = z.real();
            pt.z_im = z.imag();
            pt.s11_db = g > 1e-12 ? 20.0 * std::log10(g) : -200.0;
            pt.vswr = vswr_of(g);
        }
        pts.push_back(pt);
    }
    emfd_port_sweep ps{};
    size_t best = 0;
    for (size_t i = 1; i < pts.size(); ++i)
        if (pts[i].s11_db < pts[best].s11_db) best = i;
    ps.best = static_cast<int64_t>(best);
    double nuBest = 1.0 / pts[best].wavelength;
    if (best > 0 && best + 1 < pts.size()) {
        const double a = pts[best - 1].s11_db, b = pts[best].s11_db, c = pts[best + 1].s11_db;
        const double denom = a - 2.0 * b + c;
        const double dnu = (band.nu_hi - band.nu_lo) / (npts - 1);
        if (std::abs(denom) > 1e-12) nuBest += std::clamp(0.5 * (a - c) / denom, -1.0, 1.0) * dnu;
    }
    ps.best_wavelength = 1.0 / nuBest;
    if (pts[best].s11_db < -10.0) {
        auto edge = [&](int dir) {
            int i = static_cast<int>(best);
            while (i + dir >= 0 && i + dir < static_cast<int>(pts.size()) &&
                   pts[static_cast<size_t>(i + dir)].s11_db < -10.0)
                i += dir;
            const int j = i + dir;
            const double nuI = 1.0 / pts[static_cast<size_t>(i)].wavelength;
            if (j < 0 || j >= static_cast<int>(pts.size())) return nuI;
            const double nuJ = 1.0 / pts[static_cast<size_t>(j)].wavelength;
            const double sI = pts[static_cast<size_t>(i)].s11_db, sJ = pts[static_cast<size_t>(j)].s11_db;
            return nuI + (nuJ - nuI) * (-10.0 - sI) / (sJ - sI);
        };
        ps.has_band = 1;
        ps.band_low_wavelength = 1.0 / edge(-1);
        ps.band_high_wavelength = 1.0 / edge(+1);
    }
    if (points) std::copy(pts.begin(), pts.end(), points);
    if (summary) *summary = ps;
}

void port_rotation(float wavelength, float dt, uint64_t step_after, float rot[4]) { rotation(wavelength, dt, step_after, rot); }
float port_blend(float averaging, float wavelength, float dt) { return blend_of(averaging, wavelength, dt); }

namespace {

void build_directions(NtffState &st, const float axis[3]) {
    // FarFieldMonitor::buildDirections (src/FarField.cpp:201-257)
    st.dirs.clear();
    const float cuts[3][2][3] = {{{1, 0, 0}, {0, 1, 0}}, {{1, 0, 0}, {0, 0, 1}}, {{0, 1, 0}, {0, 0, 1}}};
    for (const auto &c : cuts)
        for (int i = 0; i < kCutSamples; ++i) {
            const float deg = 360.0f * static_cast<float>(i) / kCutSamples;
            const float rad = deg * (3.14159265358979323846f / 180.0f);
            const float cs = std::cos(rad), sn = std::sin(rad);
            for (int a = 0; a < 3; ++a) st.dirs.push_back(cs * c[0][a] + sn * c[1][a]);
            st.dirs.push_back(0.0f);
        }
    int ax = 2;
    for (int i = 0; i < 3; ++i)
        if (std::abs(axis[i]) > std::abs(axis[ax])) ax = i;
    float n[3] = {0, 0, 0}, eu[3], ev[3];
    n[ax] = axis[ax] >= 0.0f ? 1.0f : -1.0f;
    if (std::abs(n[2]) > 0.5f) {  // boresightFrame (src/FarField.hpp:29-37)
        eu[0] = 1; eu[1] = 0; eu[2] = 0;
        ev[0] = 0; ev[1] = 1; ev[2] = 0;
    } else {
        ev[0] = 0; ev[1] = 0; ev[2] = 1;
        eu[0] = n[1] * ev[2] - ev[1] * n[2];
        eu[1] = n[2] * ev[0] - ev[2] * n[0];
        eu[2] = n[0] * ev[1] - ev[0] * n[1];
    }
    for (int a = 0; a < 3; ++a) {
        st.map_normal[a] = n[a];
        st.frame[a] = n[a];
        st.frame[3 + a] = eu[a];
        st.frame[6 + a] = ev[a];
    }
    st.map_index.assign(2 * kMapSize * kMapSize, -1);
    for (int h = 0; h < 2; ++h)
        for (int r = 0; r < kMapSize; ++r)
            for (int c = 0; c < kMapSize; ++c) {
                const float u = -1.0f + 2.0f * static_cast<float>(c) / (kMapSize - 1);
                const float v = 1.0f - 2.0f * static_cast<float>(r) / (kMapSize - 1);
                const float rho2 = u * u + v * v;
                if (rho2 > 1.0f) continue;
                const float w = std::sqrt(std::max(1.0f - rho2, 0.0f));
                st.map_index[static_cast<size_t>((h * kMapSize + r) * kMapSize + c)] =
                    static_cast<int>(st.dirs.size() / 4);
                const float ww = h == 0 ? w : -w;
                for (int a = 0; a < 3; ++a) st.dirs.push_back(u * eu[a] + v * ev[a] + ww * n[a]);
                st.dirs.push_back(0.0f);
            }
}

// analyseCut (src/FarField.cpp:160-186): out = peak dBi, peak angle, beamwidth, front/back
void analyse_cut(const std::vector<float> &dbi, const std::vector<float> &angle, float out[4]) {
    const int n = static_cast<int>(dbi.size());
    const int peak = static_cast<int>(std::max_element(dbi.begin(), dbi.end()) - dbi.begin());
    const float peakDbi = dbi[peak];
    const float step = 360.0f / static_cast<float>(n);
    const float threshold = peakDbi - 3.0f;
    auto at = [&](int i) { return dbi[static_cast<size_t>(((i % n) + n) % n)]; };
    float half[2] = {-1.0f, -1.0f};
    for (int dir = 0; dir < 2; ++dir) {
        const int sgn = dir == 0 ? 1 : -1;
        for (int k = 1; k <= n / 2; ++k) {
            const float a = at(peak + sgn * (k - 1)), b = at(peak + sgn * k);
            if (b < threshold) {
                const float frac = (a - threshold) / std::max(a - b, 1e-6f);
                half[dir] = (static_cast<float>(k - 1) + frac) * step;
                break;
            }
        }
    }
    out[0] = peakDbi;
    out[1] = angle[static_cast<size_t>(peak)];
    out[2] = (half[0] < 0.0f || half[1] < 0.0f) ? 360.0f : half[0] + half[1];
    out[3] = peakDbi - at(peak + n / 2);
}

float to_dbi(double d) { return d > 1e-12 ? static_cast<float>(10.0 * std::log10(d)) : -120.0f; }

}  // namespace

// ---- hooks used by emfd_grid_step ----
int monitors_consts_floats(emfd_ctx *ctx) {
    // port rotation (4), far-field rotation (4 per wavelength), line-monitor rotation (4)
    return 4 + 4 * ctx->ntff.n_freq() + 4;
}

int monitors_fill_consts(emfd_ctx *ctx, float *row, uint64_t step_after) {
    if (ctx->ports.enabled()) rotation(ctx->ports.wavelength, ctx->cfg.courant, step_after, row);
    if (ctx->ntff.enabled)
        for (int f = 0; f < ctx->ntff.n_freq(); ++f)
            rotation(ctx->ntff.wavelength(f), ctx->cfg.courant, step_after, row + 4 + 4 * f);
    if (ctx->lines.enabled())  // src/Coax.cpp:177-180
        rotation(std::max(ctx->lines.wavelength, 2.0f), ctx->cfg.courant, step_after, row + 4 + 4 * ctx->ntff.n_freq());
    return EMFD_OK;
}

int monitors_dft_freqs(const emfd_ctx *ctx) {
    const int np = ctx->ports.enabled() ? static_cast<int>(ctx->ports.dft_wavelengths.size()) : 0;
    const int nf = ctx->ntff.enabled && ctx->ntff.dft ? ctx->ntff.n_freq() : 0;
    return np + nf;
}

// FP64 rotations of the running DFTs for the `count` steps of one step call: per step, per
// frequency (ports' first, then the far-field wavelengths) cos w tE, sin w tE, cos w tH, sin w tH
// with tE = n dt (n = the step count after the step) and tH = tE - dt / 2.
int monitors_upload_dft(emfd_ctx *ctx, int count) {
    const int nd = monitors_dft_freqs(ctx);
    if (nd == 0) return EMFD_OK;
    std::vector<double> wl;
    if (ctx->ports.enabled())
        for (double w : ctx->ports.dft_wavelengths) wl.push_back(w);
    if (ctx->ntff.enabled && ctx->ntff.dft)
        for (int f = 0; f < ctx->ntff.n_freq(); ++f) wl.push_back(static_cast<double>(ctx->ntff.wavelength(f)));
    const double dt = static_cast<double>(ctx->cfg.courant);
    std::vector<double> &h = ctx->h_consts64;
    h.assign(static_cast<size_t>(count) * nd * 4, 0.0);
    for (int k = 0; k < count; ++k)
        for (int q = 0; q < nd; ++q) {
            const double w = kTwoPi / wl[static_cast<size_t>(q)];
            const double tE = static_cast<double>(ctx->step_count + static_cast<uint64_t>(k) + 1) * dt;
            const double tH = tE - 0.5 * dt;
            double *r = &h[(static_cast<size_t>(k) * nd + q) * 4];
            r[0] = std::cos(w * tE);
            r[1] = std::sin(w * tE);
            r[2] = std::cos(w * tH);
            r[3] = std::sin(w * tH);
        }
    if (ctx->d_consts64.bytes < h.size() * sizeof(double)) {
        EMFD_HIP(hipStreamSynchronize(ctx->stream));  // earlier steps may still read the old buffer
        if (int rc = ctx->d_consts64.ensure(h.size() * sizeof(double))) return rc;
    }
    EMFD_HIP(hipMemcpyAsync(ctx->d_consts64.ptr, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream));
    return EMFD_OK;
}

int monitors_after_step(emfd_ctx *ctx, const float *consts, int stride, int k, int count) {
    const LatticeArgs lat = ctx->lattice();
    const FieldPtrs f = ctx->fields(ctx->cur);
    PortsState &P = ctx->ports;
    NtffState &N = ctx->ntff;
    PortJob pj;
    if (P.enabled()) {
        int64_t rec = -1;
        if (P.recording() && P.recorded < P.capacity_steps) rec = P.recorded++;
        pj.ports = P.d_ports.as<PortDev>();
        pj.n = static_cast<int>(P.ports.size());
        pj.rot = consts;
        pj.blend = blend_of(P.averaging, P.wavelength, ctx->cfg.courant);
        pj.phasors = P.d_phasors.as<float>();
        pj.series = P.d_series.as<float>();
        pj.rec = rec;
        if (!N.enabled || N.box.n_points <= 0) {
            EMFD_HIP(launch_port_accumulate(lat, f, pj.ports, pj.n, pj.rot, pj.blend, pj.phasors, pj.series, pj.rec,
                                            ctx->stream));
            pj = PortJob{};
        }
        if (!P.dft_wavelengths.empty()) {
            const int nd = monitors_dft_freqs(ctx);
            EMFD_HIP(launch_port_dft(lat, f, P.d_ports.as<PortDev>(), static_cast<int>(P.ports.size()),
                                     ctx->d_consts64.as<double>() + static_cast<size_t>(k) * nd * 4,
                                     static_cast<int>(P.dft_wavelengths.size()), P.d_dft.as<double>(), ctx->stream));
        }
    }
    if (N.enabled) {
        // the port step, when there is one, rides along as extra workgroups of the far-field gather
        NtffFreqs fr{};
        fr.n = N.n_freq();
        for (int q = 0; q < fr.n; ++q)
            fr.blend[q] = blend_of(N.setup.averaging_periods, N.wavelength(q), ctx->cfg.courant);
        // runs of up to kNtffRun steps inside one step call are sampled and applied together
        // (bit-identical to per-step accumulation); a call's last step always completes its run,
        // so the phasors are current whenever a call returns
        const bool last = k + 1 == count;
        if (N.pending == 0 && (last || !N.d_samples.ptr) && !N.dft) {
            EMFD_HIP(launch_ntff_accumulate(lat, f, N.box, consts + 4, fr, N.d_phasors.as<float>(), pj, ctx->stream));
        } else {
            const size_t np = static_cast<size_t>(N.box.n_points);
            EMFD_HIP(launch_ntff_sample(lat, f, N.box, N.d_samples.as<float>() + 6 * np * N.pending, pj, ctx->stream));
            ++N.pending;
            if (N.pending == kNtffRun || last) {
                const float *rot0 = consts + 4 - static_cast<ptrdiff_t>(N.pending - 1) * stride;
                EMFD_HIP(launch_ntff_flush(N.box, rot0, stride, N.pending, fr, N.d_samples.as<float>(),
                                           N.d_phasors.as<float>(), ctx->stream));
                if (N.dft) {
                    const int nd = monitors_dft_freqs(ctx);
                    const int np_dft = ctx->ports.enabled() ? static_cast<int>(ctx->ports.dft_wavelengths.size()) : 0;
                    const double *r0 = ctx->d_consts64.as<double>() +
                                       (static_cast<size_t>(k + 1 - N.pending) * nd + np_dft) * 4;
                    EMFD_HIP(launch_ntff_dft(N.box, r0, nd * 4, N.pending, N.n_freq(), N.d_samples.as<float>(),
                                             N.d_dft.as<double>(), ctx->stream));
                }
                N.pending = 0;
            }
        }
    }
    if (ctx->lines.enabled()) {
        LinesState &Ls = ctx->lines;
        // src/Coax.cpp:181: tau from the wavelength as given
        EMFD_HIP(launch_line_accumulate(lat, f, Ls.d_quant.as<int>(), Ls.n, Ls.d_terms.as<int>(),
                                        Ls.d_weights.as<float>(), consts + 4 + 4 * N.n_freq(),
                                        blend_of(Ls.averaging, Ls.wavelength, ctx->cfg.courant),
                                        Ls.d_phasors.as<float>(), Ls.d_now.as<float>(), ctx->stream));
    }
    return EMFD_OK;
}

}  // namespace emfd

using namespace emfd;

extern "C" {

// ============================================================================================
// CoaxMonitor's line monitor (src/Coax.cpp:95-198)
// ============================================================================================
int emfd_lines_configure(emfd_ctx *ctx, const int *quantities, int n, const int *terms, const float *weights,
                         int n_terms, float wavelength, float averaging_periods) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (n < 0 || n_terms < 0 || (n > 0 && !quantities) || (n_terms > 0 && (!terms || !weights)))
        return fail(EMFD_ERR_INVALID, "null or negative line-monitor arrays");
    for (int q = 0; q < n; ++q) {
        const int first = quantities[4 * q], count = quantities[4 * q + 1], kind = quantities[4 * q + 2];
        if (first < 0 || count < 0 || first > n_terms - count || (kind != 0 && kind != 1))
            return fail(EMFD_ERR_INVALID, "line-monitor quantity outside the term list or of unknown kind");
    }
    for (int k = 0; k < n_terms; ++k)
        if (terms[4 * k + 3] < 0 || terms[4 * k + 3] > 5)
            return fail(EMFD_ERR_INVALID, "line-monitor term component outside 0..5");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    LinesState &Ls = ctx->lines;
    Ls.n = 0;
    Ls.wavelength = wavelength;
    Ls.averaging = std::max(averaging_periods, 0.5f);
    Ls.start_time = ctx->time();
    if (n == 0) return EMFD_OK;
    const size_t nt = static_cast<size_t>(std::max(n_terms, 1));
    if (int rc = Ls.d_quant.ensure(static_cast<size_t>(n) * 4 * sizeof(int))) return rc;
    if (int rc = Ls.d_terms.ensure(nt * 4 * sizeof(int))) return rc;
    if (int rc = Ls.d_weights.ensure(nt * sizeof(float))) return rc;
    if (int rc = Ls.d_phasors.ensure(static_cast<size_t>(n) * 4 * sizeof(float))) return rc;
    if (int rc = Ls.d_now.ensure(static_cast<size_t>(n) * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpy(Ls.d_quant.ptr, quantities, static_cast<size_t>(n) * 4 * sizeof(int), hipMemcpyHostToDevice));
    if (n_terms > 0) {
        EMFD_HIP(hipMemcpy(Ls.d_terms.ptr, terms, static_cast<size_t>(n_terms) * 4 * sizeof(int),
                           hipMemcpyHostToDevice));
        EMFD_HIP(hipMemcpy(Ls.d_weights.ptr, weights, static_cast<size_t>(n_terms) * sizeof(float),
                           hipMemcpyHostToDevice));
    }
    EMFD_HIP(hipMemset(Ls.d_phasors.ptr, 0, static_cast<size_t>(n) * 4 * sizeof(float)));
    EMFD_HIP(hipMemset(Ls.d_now.ptr, 0, static_cast<size_t>(n) * sizeof(float)));
    Ls.n = n;
    return EMFD_OK;
}

int emfd_lines_read(emfd_ctx *ctx, float *phasors, float *now, int n) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    const LinesState &Ls = ctx->lines;
    if (n < Ls.n) return fail(EMFD_ERR_INVALID, "line-monitor output smaller than its quantity count");
    if (Ls.n == 0) return EMFD_OK;
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    if (phasors)
        EMFD_HIP(hipMemcpy(phasors, Ls.d_phasors.ptr, static_cast<size_t>(Ls.n) * 4 * sizeof(float),
                           hipMemcpyDeviceToHost));
    if (now) EMFD_HIP(hipMemcpy(now, Ls.d_now.ptr, static_cast<size_t>(Ls.n) * sizeof(float), hipMemcpyDeviceToHost));
    return EMFD_OK;
}

double emfd_coax_geometry_factor(int inner, int outer) {
    // CoaxMonitor::geometryFactor (src/Coax.cpp:42-83): Laplace's equation on the (outer + 1)^2
    // nodes of the cross-section, inner conductor nodes at 1 and the shield ring at 0, by
    // successive over-relaxation in row order with the optimal factor; G = the sum over the
    // lattice edges of the squared potential difference
    const int n = outer + 1;
    if (n < 2) return 0.0;
    const int lo = (outer - inner) / 2, hi = lo + inner;
    std::vector<double> phi(static_cast<size_t>(n) * n, 0.0);
    std::vector<char> fixed(phi.size(), 0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const size_t k = static_cast<size_t>(j) * n + i;
            if (i == 0 || j == 0 || i == n - 1 || j == n - 1) fixed[k] = 1;
            if (i >= lo && i <= hi && j >= lo && j <= hi) {
                fixed[k] = 1;
                phi[k] = 1.0;
            }
        }
    const double omega = 2.0 / (1.0 + std::sin(3.14159265358979323846 / n));
    for (int it = 0; it < 20000; ++it) {
        double change = 0.0;
        for (int j = 1; j < n - 1; ++j)
            for (int i = 1; i < n - 1; ++i) {
                const size_t k = static_cast<size_t>(j) * n + i;
                if (fixed[k]) continue;
                const double d = omega * (0.25 * (phi[k - 1] + phi[k + 1] + phi[k - n] + phi[k + n]) - phi[k]);
                phi[k] += d;
                change = std::max(change, std::abs(d));
            }
        if (change < 1e-13) break;
    }
    double g = 0.0;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const size_t k = static_cast<size_t>(j) * n + i;
            if (i + 1 < n) g += (phi[k + 1] - phi[k]) * (phi[k + 1] - phi[k]);
            if (j + 1 < n) g += (phi[k + n] - phi[k]) * (phi[k + n] - phi[k]);
        }
    return g;
}

// ============================================================================================
// PortMonitor
// ============================================================================================
int emfd_ports_configure(emfd_ctx *ctx, const emfd_port_setup *ports, int n, float averaging_periods,
                         const emfd_sweep_setup *sweep, float cell_size, int64_t record_steps) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (n > 0 && !ports) return fail(EMFD_ERR_INVALID, "null ports");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    PortsState &P = ctx->ports;
    // src/Ports.cpp:187-198
    P.ports.assign(ports, ports + std::min(std::max(n, 0), kMaxPorts));
    P.averaging = std::max(averaging_periods, 0.5f);
    P.wavelength = P.ports.empty() ? 16.0f : std::max(P.ports.front().wavelength, 2.0f);
    P.start_time = ctx->time();
    P.cell_size = cell_size;
    P.has_sweep = sweep != nullptr;
    if (sweep) P.sweep = *sweep;
    P.recorded = 0;
    P.capacity_steps = 0;
    P.record_steps = std::max<int64_t>(record_steps, 0);
    P.dft_wavelengths.clear();
    if (P.ports.empty()) return EMFD_OK;
    std::vector<PortDev> packed;
    for (const emfd_port_setup &p : P.ports) {
        PortDev d{};
        for (int a = 0; a < 3; ++a) d.texel[a] = p.texel[a];
        d.axis = std::clamp(p.axis, 0, 2);
        d.edges = p.edges;
        d.sign = p.sign >= 0.0f ? 1 : -1;
        packed.push_back(d);
    }
    if (int rc = P.d_ports.ensure(packed.size() * sizeof(PortDev))) return rc;
    if (int rc = P.d_phasors.ensure(packed.size() * 8 * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpy(P.d_ports.ptr, packed.data(), packed.size() * sizeof(PortDev), hipMemcpyHostToDevice));
    EMFD_HIP(hipMemset(P.d_phasors.ptr, 0, packed.size() * 8 * sizeof(float)));
    if (P.recording()) {
        P.capacity_steps = static_cast<int64_t>(kSeriesFloats / (2 * P.ports.size()));
        if (!P.has_sweep) P.capacity_steps = std::min(P.capacity_steps, P.record_steps);
        if (int rc = P.d_series.ensure(static_cast<size_t>(P.capacity_steps) * P.ports.size() * 2 * sizeof(float)))
            return rc;
    }
    return EMFD_OK;
}

int emfd_ports_disable(emfd_ctx *ctx) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->ports.ports.clear();
    ctx->ports.has_sweep = false;
    ctx->ports.recorded = 0;
    return EMFD_OK;
}

int emfd_ports_set_dft(emfd_ctx *ctx, const double *wavelengths, int n) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    PortsState &P = ctx->ports;
    if (n < 0 || n > 4096 || (n > 0 && !wavelengths)) return fail(EMFD_ERR_INVALID, "0 to 4096 wavelengths");
    if (n > 0 && !P.enabled()) return fail(EMFD_ERR_STATE, "port monitor not configured");
    for (int q = 0; q < n; ++q)
        if (!(wavelengths[q] > 0.0)) return fail(EMFD_ERR_INVALID, "running-DFT wavelengths must be positive");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    P.dft_wavelengths.assign(wavelengths, wavelengths + n);
    if (n == 0) return EMFD_OK;
    const size_t dbl = P.ports.size() * static_cast<size_t>(n) * 4;
    if (int rc = P.d_dft.ensure(dbl * sizeof(double))) return rc;
    EMFD_HIP(launch_fill(P.d_dft.as<float>(), 2 * dbl, 0.0f, ctx->stream));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    return EMFD_OK;
}

int emfd_ports_read_dft(emfd_ctx *ctx, double *out, int64_t n_values) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    PortsState &P = ctx->ports;
    if (P.dft_wavelengths.empty()) return fail(EMFD_ERR_STATE, "port running DFT not enabled");
    const size_t dbl = P.ports.size() * P.dft_wavelengths.size() * 4;
    if (n_values < static_cast<int64_t>(dbl)) return fail(EMFD_ERR_INVALID, "output too small");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpy(out, P.d_dft.ptr, dbl * sizeof(double), hipMemcpyDeviceToHost));
    return EMFD_OK;
}

int emfd_ports_read_phasors(emfd_ctx *ctx, float *out, int n) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    const int np = static_cast<int>(ctx->ports.ports.size());
    if (n < np) return fail(EMFD_ERR_INVALID, "output too small");
    if (np == 0) return EMFD_OK;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpy(out, ctx->ports.d_phasors.ptr, static_cast<size_t>(np) * 8 * sizeof(float),
                       hipMemcpyDeviceToHost));
    return EMFD_OK;
}

int emfd_ports_evaluate(emfd_ctx *ctx, emfd_port_reading *out, int n) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    const int np = static_cast<int>(ctx->ports.ports.size());
    if (n < np) return fail(EMFD_ERR_INVALID, "output too small");
    std::vector<float> ph(static_cast<size_t>(np) * 8);
    if (np > 0)
        if (int rc = emfd_ports_read_phasors(ctx, ph.data(), np)) return rc;
    // PortMonitor::computeReadings (src/Ports.cpp:159-185)
    for (int i = 0; i < np; ++i) {
        const emfd_port_setup &p = ctx->ports.ports[i];
        const float *s2 = &ph[8 * i + 4];
        emfd_port_reading r{};
        const std::complex<double> V = -std::complex<double>(s2[0], s2[1]);
        const std::complex<double> I(s2[2], s2[3]);
        r.v_re = V.real(); r.v_im = V.imag();
        r.i_re = I.real(); r.i_im = I.imag();
        r.valid = std::abs(I) > 1e-12 && std::isfinite(std::abs(V));
        if (r.valid) {
            const std::complex<double> z = V / I * kEta0;
            const double z0 = std::max(static_cast<double>(p.impedance_ohm), 1e-6);
            const std::complex<double> g = reflection(z, z0);
            const double gm = std::abs(g);
            r.z_re = z.real(); r.z_im = z.imag();
            r.gamma_re = g.real(); r.gamma_im = g.imag();
            r.vswr = vswr_of(gm);
            r.return_loss_db = gm > 1e-12 ? -20.0 * std::log10(gm) : 200.0;
            r.accepted_power = 0.5 * std::real(V * std::conj(I));
        }
        if (p.voltage_feed) {
            const double rs = std::max(static_cast<double>(p.impedance_ohm), 1e-6) / kEta0;
            r.available_power = static_cast<double>(p.drive) * p.drive / (8.0 * rs);
        }
        out[i] = r;
    }
    return EMFD_OK;
}

int64_t emfd_ports_recorded_steps(emfd_ctx *ctx) { return ctx ? ctx->ports.recorded : -1; }

int emfd_ports_read_series(emfd_ctx *ctx, float *out, int64_t cap_floats) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    const size_t need = static_cast<size_t>(ctx->ports.recorded) * ctx->ports.ports.size() * 2;
    if (static_cast<int64_t>(need) > cap_floats) return fail(EMFD_ERR_INVALID, "output too small");
    if (need == 0) return EMFD_OK;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpy(out, ctx->ports.d_series.ptr, need * sizeof(float), hipMemcpyDeviceToHost));
    return EMFD_OK;
}

int emfd_ports_evaluate_sweep(emfd_ctx *ctx, double dt, emfd_sweep_point *points, emfd_port_sweep *summary,
                              double *residual, int64_t *samples, int *full) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    PortsState &P = ctx->ports;
    if (!P.has_sweep || P.ports.empty() || P.recorded < 16)
        return fail(EMFD_ERR_STATE, "no sweep configured or fewer than 16 recorded steps");
    const size_t nPorts = P.ports.size();
    const size_t n = static_cast<size_t>(P.recorded);
    std::vector<float> series(n * nPorts * 2);
    if (int rc = emfd_ports_read_series(ctx, series.data(), static_cast<int64_t>(series.size()))) return rc;
    // PortMonitor::evaluateSweep (src/Ports.cpp:187-278)
    const SweepBand band = sweep_band(P.sweep);
    double peak = 0.0, late = 0.0;
    for (size_t k = 0; k < n; ++k) {
        const double v = std::abs(static_cast<double>(series[2 * (k * nPorts)]));
        peak = std::max(peak, v);
        if (k >= n - std::max<size_t>(n / 20, 1)) late = std::max(late, v);
    }
    if (residual) *residual = peak > 0.0 ? late / peak : 1.0;
    if (samples) *samples = static_cast<int64_t>(n);
    if (full) *full = P.recorded >= P.capacity_steps;
    std::vector<std::complex<double>> sumV(static_cast<size_t>(band.points)), sumI(static_cast<size_t>(band.points));
    for (size_t p = 0; p < nPorts; ++p) {
        for (int f = 0; f < band.points; ++f) {
            const double w = kTwoPi * band.nu(f);
            const std::complex<double> step = std::polar(1.0, -w * dt);
            std::complex<double> rot(1.0, 0.0), sv(0.0), si(0.0);
            for (size_t k = 0; k < n; ++k) {
                sv += static_cast<double>(series[2 * (k * nPorts + p)]) * rot;
                si += static_cast<double>(series[2 * (k * nPorts + p) + 1]) * rot;
                rot *= step;
                if ((k & 1023) == 1023) rot /= std::abs(rot);
            }
            sumV[static_cast<size_t>(f)] = sv;
            sumI[static_cast<size_t>(f)] = si * std::polar(1.0, w * 0.5 * dt);
        }
        sweep_finish(band, sumV.data(), sumI.data(), static_cast<double>(P.ports[p].impedance_ohm),
                     points ? points + p * static_cast<size_t>(band.points) : nullptr, summary ? summary + p : nullptr);
    }
    return EMFD_OK;
}

// ============================================================================================
// FarFieldMonitor
// ============================================================================================
int emfd_ntff_configure(emfd_ctx *ctx, const emfd_far_field_setup *setup, emfd_far_field_setup *applied,
                        int *n_points) {
    if (!ctx || !setup) return fail(EMFD_ERR_INVALID, "null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    NtffState &N = ctx->ntff;
    // src/FarField.cpp:259-307
    const int L = ctx->cfg.pml_thickness;
    emfd_far_field_setup s = *setup;
    for (int a = 0; a < 3; ++a) {
        const int lo = L + 1, hi = ctx->cfg.size[a] - L - 2;
        const int mn = std::min(setup->box_min[a], setup->box_max[a]);
        const int mx = std::max(setup->box_min[a], setup->box_max[a]);
        s.box_min[a] = std::clamp(mn, lo, std::max(lo, hi));
        s.box_max[a] = std::clamp(mx, lo, std::max(lo, hi));
        // glm::clamp(v, lo, hi) with hi < lo is undefined in the reference; lo wins here
    }
    for (int a = 0; a < 3; ++a)
        if (s.box_max[a] - s.box_min[a] < 4)
            return fail(EMFD_ERR_INVALID,
                        "far_field box is degenerate after clamping to the region inside the absorbing layer");
    s.wavelength = std::max(s.wavelength, 2.0f);
    const float dx = s.map_axis[0] - N.map_normal[0], dy = s.map_axis[1] - N.map_normal[1],
                dz = s.map_axis[2] - N.map_normal[2];
    if (std::sqrt(dx * dx + dy * dy + dz * dz) > 1e-3f || N.dirs.empty()) build_directions(N, s.map_axis);
    s.averaging_periods = std::max(s.averaging_periods, 0.5f);
    NtffBox b{};
    b.face_offset[0] = 0;
    for (int f = 0; f < 6; ++f) {
        const int a = f >> 1, u = a == 0 ? 1 : 0, v = a == 2 ? 1 : 2;
        b.face_offset[f + 1] = b.face_offset[f] + (s.box_max[u] - s.box_min[u] + 1) * (s.box_max[v] - s.box_min[v] + 1);
    }
    for (int a = 0; a < 3; ++a) {
        b.bmin[a] = s.box_min[a];
        b.bmax[a] = s.box_max[a];
    }
    b.n_points = b.face_offset[6];
    N.wavelengths.assign(1, s.wavelength);  // configure resets the list to the setup's wavelength
    if (int rc = N.d_phasors.ensure(static_cast<size_t>(b.n_points) * 24 * sizeof(float))) return rc;
    if (int rc = N.d_samples.ensure(static_cast<size_t>(b.n_points) * 6 * kNtffRun * sizeof(float))) return rc;
    N.pending = 0;
    // side columns x = X-1, X of the two x faces over y, z in [box min - 1, box max]
    N.side = SideCols{};
    N.side.x[0] = b.bmin[0];
    N.side.x[1] = b.bmax[0];
    N.side.y0 = b.bmin[1] - 1;
    N.side.ny = b.bmax[1] - b.bmin[1] + 2;
    N.side.z0 = b.bmin[2] - 1;
    N.side.nz = b.bmax[2] - b.bmin[2] + 2;
    N.side.gy0 = N.side.y0 / kSideGroup;
    N.side.nyp = ((N.side.y0 + N.side.ny - 1) / kSideGroup - N.side.gy0 + 1) * 8;
    if (int rc = N.d_side.ensure(static_cast<size_t>(2 * 6 * 2) * N.side.nyp * N.side.nz * sizeof(float))) return rc;
    EMFD_HIP(hipMemset(N.d_phasors.ptr, 0, static_cast<size_t>(b.n_points) * 24 * sizeof(float)));
    N.box = b;
    N.ff = ff_geometry(b, &N.ff_halves);
    N.ff_zeroed = false;
    N.dft = false;
    N.setup = s;
    N.enabled = true;
    N.start_time = ctx->time();
    if (applied) *applied = s;
    if (n_points) *n_points = b.n_points;
    return EMFD_OK;
}

int emfd_ntff_disable(emfd_ctx *ctx) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    ctx->ntff.enabled = false;
    return EMFD_OK;
}

int emfd_ntff_reset(emfd_ctx *ctx) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    if (!ctx->ntff.enabled) return EMFD_OK;
    EMFD_HIP(launch_fill(ctx->ntff.d_phasors.as<float>(),
                         static_cast<size_t>(ctx->ntff.box.n_points) * 24 * ctx->ntff.n_freq(), 0.0f, ctx->stream));
    return EMFD_OK;
}

int emfd_ntff_set_wavelengths(emfd_ctx *ctx, const float *wavelengths, int n) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    NtffState &N = ctx->ntff;
    if (!N.enabled) return fail(EMFD_ERR_STATE, "far-field monitor not configured");
    if (n < 1 || n > kMaxNtffFreq || !wavelengths) return fail(EMFD_ERR_INVALID, "1 to 8 wavelengths");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    N.wavelengths.clear();
    for (int f = 0; f < n; ++f) N.wavelengths.push_back(std::max(wavelengths[f], 2.0f));  // as configure
    N.setup.wavelength = N.wavelengths[0];
    const size_t floats = static_cast<size_t>(N.box.n_points) * 24 * n;
    if (int rc = N.d_phasors.ensure(floats * sizeof(float))) return rc;
    EMFD_HIP(launch_fill(N.d_phasors.as<float>(), floats, 0.0f, ctx->stream));
    N.start_time = ctx->time();
    if (N.dft) {
        // the running DFTs follow the wavelength list, restarted
        const size_t dbl = static_cast<size_t>(N.box.n_points) * 12 * n;
        if (int rc = N.d_dft.ensure(dbl * sizeof(double))) return rc;
        EMFD_HIP(launch_fill(N.d_dft.as<float>(), 2 * dbl, 0.0f, ctx->stream));
    }
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    return EMFD_OK;
}

int emfd_ntff_set_dft(emfd_ctx *ctx, int enable) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    NtffState &N = ctx->ntff;
    if (!N.enabled) return fail(EMFD_ERR_STATE, "far-field monitor not configured");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    N.dft = enable != 0;
    if (!N.dft) return EMFD_OK;
    const size_t dbl = static_cast<size_t>(N.box.n_points) * 12 * N.n_freq();
    if (int rc = N.d_dft.ensure(dbl * sizeof(double))) return rc;
    EMFD_HIP(launch_fill(N.d_dft.as<float>(), 2 * dbl, 0.0f, ctx->stream));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    return EMFD_OK;
}

int emfd_ntff_read_dft(emfd_ctx *ctx, int freq, double *out, int64_t n_points) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    NtffState &N = ctx->ntff;
    if (!N.enabled || !N.dft) return fail(EMFD_ERR_STATE, "far-field running DFT not enabled");
    if (freq < 0 || freq >= N.n_freq()) return fail(EMFD_ERR_INVALID, "no such far-field wavelength");
    const size_t np = static_cast<size_t>(N.box.n_points);
    if (n_points < static_cast<int64_t>(np)) return fail(EMFD_ERR_INVALID, "output too small");
    EMFD_HIP(hipSetDevice(ctx->device));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<double> soa(np * 12);
    EMFD_HIP(hipMemcpy(soa.data(), N.d_dft.as<double>() + np * 12 * static_cast<size_t>(freq),
                       soa.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < np; ++i)
        for (size_t j = 0; j < 12; ++j) out[i * 12 + j] = soa[j * np + i];
    return EMFD_OK;
}

int emfd_ntff_read_phasors(emfd_ctx *ctx, float *out, int64_t n_points) {
    return emfd_ntff_read_phasors_at(ctx, 0, out, n_points);
}

int emfd_ntff_read_phasors_at(emfd_ctx *ctx, int freq, float *out, int64_t n_points) {
    if (!ctx || !out) return fail(EMFD_ERR_INVALID, "null argument");
    if (!ctx->ntff.enabled) return fail(EMFD_ERR_STATE, "far-field monitor not configured");
    if (freq < 0 || freq >= ctx->ntff.n_freq()) return fail(EMFD_ERR_INVALID, "no such far-field wavelength");
    const size_t np = static_cast<size_t>(ctx->ntff.box.n_points);
    if (n_points < static_cast<int64_t>(np)) return fail(EMFD_ERR_INVALID, "output too small");
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> soa(np * 24);
    EMFD_HIP(hipMemcpy(soa.data(), ctx->ntff.d_phasors.as<float>() + np * 24 * static_cast<size_t>(freq),
                       soa.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < np; ++i) {
        float *o = out + 32 * i;
        for (int s = 0; s < 2; ++s)
            for (int q = 0; q < 4; ++q) {
                for (int c = 0; c < 3; ++c) o[16 * s + 4 * q + c] = soa[static_cast<size_t>(12 * s + 3 * q + c) * np + i];
                o[16 * s + 4 * q + 3] = 0.0f;
            }
    }
    return EMFD_OK;
}

int emfd_ntff_directions(emfd_ctx *ctx, float *dirs, int cap, int *n_dirs, int *map_index, float *frame) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    NtffState &N = ctx->ntff;
    if (N.dirs.empty()) {
        const float z[3] = {0, 0, 1};
        build_directions(N, z);
    }
    const int nd = static_cast<int>(N.dirs.size() / 4);
    if (n_dirs) *n_dirs = nd;
    if (dirs) {
        if (cap < nd) return fail(EMFD_ERR_INVALID, "direction buffer too small");
        std::copy(N.dirs.begin(), N.dirs.end(), dirs);
    }
    if (map_index) std::copy(N.map_index.begin(), N.map_index.end(), map_index);
    if (frame) std::copy(N.frame, N.frame + 9, frame);
    return EMFD_OK;
}

int emfd_ntff_eval(emfd_ctx *ctx, const float *dirs, int n_dirs, float *f2, double *power) {
    return emfd_ntff_eval_at(ctx, 0, dirs, n_dirs, f2, power);
}

namespace {
int ntff_eval_impl(emfd_ctx *ctx, int freq, bool from_dft, const float *dirs, int n_dirs, float *f2, double *power);
}

int emfd_ntff_eval_at(emfd_ctx *ctx, int freq, const float *dirs, int n_dirs, float *f2, double *power) {
    return ntff_eval_impl(ctx, freq, false, dirs, n_dirs, f2, power);
}

int emfd_ntff_eval_dft(emfd_ctx *ctx, int freq, const float *dirs, int n_dirs, float *f2, double *power) {
    return ntff_eval_impl(ctx, freq, true, dirs, n_dirs, f2, power);
}

namespace {
int ntff_eval_impl(emfd_ctx *ctx, int freq, bool from_dft, const float *dirs, int n_dirs, float *f2, double *power) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    NtffState &N = ctx->ntff;
    if (!N.enabled) return fail(EMFD_ERR_STATE, "far-field monitor not configured");
    if (from_dft && !N.dft) return fail(EMFD_ERR_STATE, "far-field running DFT not enabled");
    if (freq < 0 || freq >= N.n_freq()) return fail(EMFD_ERR_INVALID, "no such far-field wavelength");
    if (n_dirs < 0 || (n_dirs > 0 && (!dirs || !f2))) return fail(EMFD_ERR_INVALID, "bad direction arguments");
    EMFD_HIP(hipSetDevice(ctx->device));
    const int pgroups = (N.box.n_points + 255) / 256;
    if (int rc = N.d_dirs.ensure(static_cast<size_t>(std::max(n_dirs, 1)) * 4 * sizeof(float))) return rc;
    if (int rc = N.d_partials.ensure(static_cast<size_t>(std::max(n_dirs, 1)) * 6 * 12 * sizeof(float))) return rc;
    if (int rc = N.d_out.ensure(static_cast<size_t>(std::max(n_dirs, 1)) * sizeof(float))) return rc;
    if (int rc = N.d_power.ensure(static_cast<size_t>(pgroups) * sizeof(double))) return rc;
    if (int rc = N.d_absmax.ensure(static_cast<size_t>(ff_absmax_blocks(N.box.n_points)) * sizeof(float))) return rc;
    if (!N.ff_zeroed) {
        // operand buffers of the box's geometry; their padding must stay zero
        if (int rc = N.d_ahi.ensure(std::max<size_t>(N.ff_halves, 8) * 2)) return rc;
        if (int rc = N.d_alo.ensure(std::max<size_t>(N.ff_halves, 8) * 2)) return rc;
        EMFD_HIP(launch_fill(N.d_ahi.as<float>(), N.ff_halves / 2, 0.0f, ctx->stream));
        EMFD_HIP(launch_fill(N.d_alo.as<float>(), N.ff_halves / 2, 0.0f, ctx->stream));
        N.ff_zeroed = true;
    }
    if (n_dirs > 0)
        EMFD_HIP(hipMemcpyAsync(N.d_dirs.ptr, dirs, static_cast<size_t>(n_dirs) * 4 * sizeof(float),
                                hipMemcpyHostToDevice, ctx->stream));
    const float k = static_cast<float>(kTwoPi / N.wavelength(freq));  // src/FarField.cpp:293
    const float *ph = N.d_phasors.as<float>() + static_cast<size_t>(N.box.n_points) * 24 * static_cast<size_t>(freq);
    if (from_dft) {
        // the FP64 sums as FP32 stage-2 planes (the pattern's inputs); |F|^2 and P scale alike
        const size_t np = static_cast<size_t>(N.box.n_points);
        if (int rc = N.d_dft_ph.ensure(np * 24 * sizeof(float))) return rc;
        EMFD_HIP(launch_ntff_dft_planes(N.d_dft.as<double>() + np * 12 * static_cast<size_t>(freq), N.box.n_points,
                                        N.d_dft_ph.as<float>(), ctx->stream));
        ph = N.d_dft_ph.as<float>();
    }
    EMFD_HIP(launch_far_field(N.box, N.ff, ph, N.d_dirs.as<float>(), n_dirs, k, N.d_absmax.as<float>(),
                              N.d_ahi.as<_Float16>(), N.d_alo.as<_Float16>(), N.d_partials.as<float>(),
                              N.d_out.as<float>(), power ? N.d_power.as<double>() : nullptr, ctx->stream));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    if (n_dirs > 0)
        EMFD_HIP(hipMemcpy(f2, N.d_out.ptr, static_cast<size_t>(n_dirs) * sizeof(float), hipMemcpyDeviceToHost));
    if (power) {
        std::vector<double> pp(static_cast<size_t>(pgroups));
        EMFD_HIP(hipMemcpy(pp.data(), N.d_power.ptr, pp.size() * sizeof(double), hipMemcpyDeviceToHost));
        double P = 0.0;
        for (double v : pp) P += v;
        *power = 0.5 * P;
    }
    return EMFD_OK;
}
}  // namespace

int emfd_ntff_evaluate(emfd_ctx *ctx, int *valid, double *radiated_power, float *max_directivity_dbi,
                       float *max_direction, float *directivity_dbi, float *cuts) {
    if (!ctx) return fail(EMFD_ERR_INVALID, "null context");
    NtffState &N = ctx->ntff;
    if (!N.enabled) return fail(EMFD_ERR_STATE, "far-field monitor not configured");
    const int nd = static_cast<int>(N.dirs.size() / 4);
    std::vector<float> f2(static_cast<size_t>(nd));
    double P = 0.0;
    if (int rc = emfd_ntff_eval(ctx, N.dirs.data(), nd, f2.data(), &P)) return rc;
    // FarFieldMonitor::finishEvaluation (src/FarField.cpp:294-335)
    const bool ok = std::isfinite(P) && P > 1e-12;
    if (valid) *valid = ok ? 1 : 0;
    if (radiated_power) *radiated_power = P;
    if (!ok) return EMFD_OK;
    const double k = kTwoPi / N.setup.wavelength;
    const double scale = k * k / (8.0 * 3.14159265358979323846 * P);
    double best = -1.0;
    int bi = 0;
    for (int i = 0; i < nd; ++i) {
        const double d = scale * static_cast<double>(f2[static_cast<size_t>(i)]);
        if (d > best) {
            best = d;
            bi = i;
        }
        if (directivity_dbi) directivity_dbi[i] = to_dbi(d);
    }
    if (max_directivity_dbi) *max_directivity_dbi = to_dbi(best);
    if (max_direction)
        for (int a = 0; a < 3; ++a) max_direction[a] = N.dirs[4 * static_cast<size_t>(bi) + a];
    if (cuts) {
        std::vector<float> angle(kCutSamples);
        for (int i = 0; i < kCutSamples; ++i) angle[i] = 360.0f * static_cast<float>(i) / kCutSamples;
        for (int c = 0; c < 3; ++c) {
            std::vector<float> dbi(kCutSamples);
            for (int i = 0; i < kCutSamples; ++i)
                dbi[i] = to_dbi(scale * static_cast<double>(f2[static_cast<size_t>(c * kCutSamples + i)]));
            float o[4];
            analyse_cut(dbi, angle, o);
            for (int j = 0; j < 4; ++j) cuts[6 * c + j] = o[j];
            cuts[6 * c + 4] = cuts[6 * c + 5] = 0.0f;
        }
    }
    return EMFD_OK;
}

// ============================================================================================
// FieldAnalyzer
// ============================================================================================
}  // extern "C"

namespace {
// The plane list of n flux surfaces with outward signs (src/FieldAnalyzer.cpp:305-338, 491-505)
// into caller-sized arrays (at most 6 planes per surface): planes clipped to the lattice, each with
// one partial per row of its rectangle.
void flux_planes(const emfd_ctx *ctx, const emfd_flux_surface *surfaces, int n, FluxPlane *planes,
                 FluxRef *refs, int *n_planes, int *total_rows, int *max_rows) {
    const int size[3] = {ctx->cfg.size[0], ctx->cfg.size[1], ctx->cfg.size[2]};
    int np = 0, total = 0, max_groups = 0;
    auto add_plane = [&](int surf, float sign, int axis, int index, int mnu, int mnv, int mxu, int mxv) {
        if (index < 0 || index >= size[axis]) return;
        const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
        mnu = std::max(mnu, 0);
        mnv = std::max(mnv, 0);
        mxu = std::min(mxu, size[u]);
        mxv = std::min(mxv, size[v]);
        if (mxu - mnu <= 0 || mxv - mnv <= 0) return;
        FluxPlane p{};
        p.axis = axis;
        p.index = index;
        p.rmin[0] = mnu;
        p.rmin[1] = mnv;
        p.rsize[0] = mxu - mnu;
        p.rsize[1] = mxv - mnv;
        p.first_group = total;
        total += p.rsize[1];
        max_groups = std::max(max_groups, p.rsize[1]);
        planes[np] = p;
        refs[np] = FluxRef{surf, sign};
        ++np;
    };
    for (int s = 0; s < n; ++s) {
        const emfd_flux_surface &fs = surfaces[s];
        if (fs.kind == 0) {
            add_plane(s, 1.0f, std::clamp(fs.axis, 0, 2), fs.index, fs.rect_min[0], fs.rect_min[1], fs.rect_max[0],
                      fs.rect_max[1]);
            continue;
        }
        int mn[3], mx[3];
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(fs.box_min[a], fs.box_max[a]);
            mx[a] = std::max(fs.box_min[a], fs.box_max[a]);
        }
        for (int axis = 0; axis < 3; ++axis) {
            const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
            add_plane(s, -1.0f, axis, mn[axis], mn[u], mn[v], mx[u] + 1, mx[v] + 1);
            add_plane(s, +1.0f, axis, mx[axis], mn[u], mn[v], mx[u] + 1, mx[v] + 1);
        }
    }
    *n_planes = np;
    *total_rows = total;
    *max_rows = max_groups;
}
}  // namespace

extern "C" {

int emfd_flux_measure(emfd_ctx *ctx, const emfd_flux_surface *surfaces, int n, double *out) {
    if (!ctx || (n > 0 && (!surfaces || !out))) return fail(EMFD_ERR_INVALID, "null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    EMFD_HIP(hipSetDevice(ctx->device));
    std::vector<FluxPlane> planes(6 * static_cast<size_t>(std::max(n, 0)));
    std::vector<FluxRef> refs(planes.size());
    int np = 0, total = 0, max_rows = 0;
    flux_planes(ctx, surfaces, n, planes.data(), refs.data(), &np, &total, &max_rows);
    for (int s = 0; s < n; ++s) out[s] = 0.0;
    if (np == 0) return EMFD_OK;
    if (int rc = ctx->scratch_misc.ensure(static_cast<size_t>(np) * sizeof(FluxPlane))) return rc;
    if (int rc = ctx->scratch_partials.ensure(static_cast<size_t>(total) * sizeof(float))) return rc;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpy(ctx->scratch_misc.ptr, planes.data(), static_cast<size_t>(np) * sizeof(FluxPlane),
                       hipMemcpyHostToDevice));
    EMFD_HIP(launch_flux(ctx->lattice(), ctx->fields(ctx->cur), ctx->scratch_misc.as<FluxPlane>(), np, max_rows,
                         ctx->scratch_partials.as<float>(), ctx->stream));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> partials(static_cast<size_t>(total));
    EMFD_HIP(hipMemcpy(partials.data(), ctx->scratch_partials.ptr, partials.size() * sizeof(float),
                       hipMemcpyDeviceToHost));
    // per plane an FP64 sum over its rows, then signed into its surface, in plane order
    for (int i = 0; i < np; ++i) {
        double sum = 0.0;
        for (int k = 0; k < planes[i].rsize[1]; ++k)
            sum += static_cast<double>(partials[static_cast<size_t>(planes[i].first_group + k)]);
        out[refs[i].surface] += static_cast<double>(refs[i].sign) * sum;
    }
    return EMFD_OK;
}

int emfd_flux_measure_noalloc(emfd_ctx *ctx, const emfd_flux_surface *surfaces, int n, double *out) {
    if (!ctx || (n > 0 && (!surfaces || !out))) return fail(EMFD_ERR_INVALID, "null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (n < 0 || n > kFluxScratchSurfaces)
        return fail(EMFD_ERR_INVALID, "emfd_flux_measure_noalloc: 0 to 64 surfaces");
    EMFD_HIP(hipSetDevice(ctx->device));
    FluxScratch sc{};
    if (int rc = flux_scratch(ctx->device, &sc)) return rc;
    FluxPlane planes[6 * kFluxScratchSurfaces];
    FluxRef refs[6 * kFluxScratchSurfaces];
    int np = 0, total = 0, max_rows = 0;
    flux_planes(ctx, surfaces, n, planes, refs, &np, &total, &max_rows);
    if (total > kFluxScratchRows)
        return fail(EMFD_ERR_INVALID, "emfd_flux_measure_noalloc: more than 262144 plane rows in all");
    for (int s = 0; s < n; ++s) out[s] = 0.0;
    if (np == 0) return EMFD_OK;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpyAsync(sc.planes, planes, static_cast<size_t>(np) * sizeof(FluxPlane), hipMemcpyHostToDevice,
                            ctx->stream));
    EMFD_HIP(hipMemcpyAsync(sc.refs, refs, static_cast<size_t>(np) * sizeof(FluxRef), hipMemcpyHostToDevice,
                            ctx->stream));
    EMFD_HIP(launch_flux(ctx->lattice(), ctx->fields(ctx->cur), sc.planes, np, max_rows, sc.partials, ctx->stream));
    EMFD_HIP(launch_flux_surfaces(sc.planes, sc.refs, np, sc.partials, n, sc.sums, ctx->stream));
    EMFD_HIP(hipMemcpyAsync(out, sc.sums, static_cast<size_t>(n) * sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    // the host arrays above stay alive until the copies have completed
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    return EMFD_OK;
}

namespace {
// sampleBlock (src/FieldAnalyzer.cpp:149-170)
float sample_block(const float *blk, const int origin[3], const int extent[3], const float q[3], int ch) {
    const float fl[3] = {std::floor(q[0]), std::floor(q[1]), std::floor(q[2])};
    const float f[3] = {q[0] - fl[0], q[1] - fl[1], q[2] - fl[2]};
    const int i0[3] = {static_cast<int>(fl[0]) - origin[0], static_cast<int>(fl[1]) - origin[1],
                       static_cast<int>(fl[2]) - origin[2]};
    float result = 0.0f;
    for (int dz = 0; dz <= 1; ++dz)
        for (int dy = 0; dy <= 1; ++dy)
            for (int dx = 0; dx <= 1; ++dx) {
                const float w = (dx ? f[0] : 1.0f - f[0]) * (dy ? f[1] : 1.0f - f[1]) * (dz ? f[2] : 1.0f - f[2]);
                if (w != 0.0f) {
                    const int x = std::clamp(i0[0] + dx, 0, extent[0] - 1), y = std::clamp(i0[1] + dy, 0, extent[1] - 1),
                              z = std::clamp(i0[2] + dz, 0, extent[2] - 1);
                    result += w * blk[4 * ((z * 3 + y) * 3 + x) + ch];
                }
            }
    return result;
}
}  // namespace

int emfd_probe_measure(emfd_ctx *ctx, const float *positions, int n, emfd_probe_sample *out) {
    if (!ctx || (n > 0 && (!positions || !out))) return fail(EMFD_ERR_INVALID, "null argument");
    if (!ctx->configured) return fail(EMFD_ERR_STATE, "grid not configured");
    if (n <= 0) return EMFD_OK;
    EMFD_HIP(hipSetDevice(ctx->device));
    const int size[3] = {ctx->cfg.size[0], ctx->cfg.size[1], ctx->cfg.size[2]};
    std::vector<int> origins(static_cast<size_t>(n) * 4);
    std::vector<float> pos(static_cast<size_t>(n) * 3);
    std::vector<int> ext(static_cast<size_t>(n) * 3);
    for (int i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const float hi = static_cast<float>(size[a] - 1);
            const float p = std::clamp(positions[3 * i + a], 0.0f, hi);
            pos[3 * i + a] = p;
            const int o = std::clamp(static_cast<int>(std::floor(p)) - 1, 0, std::max(size[a] - 3, 0));
            origins[4 * i + a] = o;
            ext[3 * i + a] = std::min(3, size[a] - o);
        }
    const size_t blk_floats = static_cast<size_t>(n) * 54 * 4;
    if (int rc = ctx->scratch_misc.ensure(origins.size() * sizeof(int))) return rc;
    if (int rc = ctx->scratch_partials.ensure(blk_floats * sizeof(float))) return rc;
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    EMFD_HIP(hipMemcpy(ctx->scratch_misc.ptr, origins.data(), origins.size() * sizeof(int), hipMemcpyHostToDevice));
    EMFD_HIP(launch_probe_gather(ctx->lattice(), ctx->fields(ctx->cur), ctx->scratch_misc.as<int>(), n,
                                 ctx->sigma_cell.as<float>(), ctx->eps_cell.as<float>(),
                                 ctx->scratch_partials.as<float>(), ctx->stream));
    EMFD_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> blocks(blk_floats);
    EMFD_HIP(hipMemcpy(blocks.data(), ctx->scratch_partials.ptr, blk_floats * sizeof(float), hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i) {
        const float *eb = &blocks[static_cast<size_t>(i) * 54 * 4];
        const float *bb = eb + 27 * 4;
        const int *o = &origins[4 * static_cast<size_t>(i)];
        const int *e = &ext[3 * static_cast<size_t>(i)];
        const float *p = &pos[3 * static_cast<size_t>(i)];
        emfd_probe_sample s{};
        const float qE[3][3] = {{p[0] - 0.5f, p[1] - 0.0f, p[2] - 0.0f},
                                {p[0] - 0.0f, p[1] - 0.5f, p[2] - 0.0f},
                                {p[0] - 0.0f, p[1] - 0.0f, p[2] - 0.5f}};
        const float qB[3][3] = {{p[0] - 0.0f, p[1] - 0.5f, p[2] - 0.5f},
                                {p[0] - 0.5f, p[1] - 0.0f, p[2] - 0.5f},
                                {p[0] - 0.5f, p[1] - 0.5f, p[2] - 0.0f}};
        for (int c = 0; c < 3; ++c) {
            s.position[c] = p[c];
            s.E[c] = sample_block(eb, o, e, qE[c], c);
            s.B[c] = sample_block(bb, o, e, qB[c], c);
        }
        s.S[0] = s.E[1] * s.B[2] - s.B[1] * s.E[2];
        s.S[1] = s.E[2] * s.B[0] - s.B[2] * s.E[0];
        s.S[2] = s.E[0] * s.B[1] - s.B[0] * s.E[1];
        int cell[3];
        for (int a = 0; a < 3; ++a) cell[a] = std::clamp(static_cast<int>(std::floor(p[a])) - o[a], 0, e[a] - 1);
        const int ci = (cell[2] * 3 + cell[1]) * 3 + cell[0];
        s.sigma = eb[4 * ci + 3];
        s.epsilon_r = bb[4 * ci + 3];
        out[i] = s;
    }
    return EMFD_OK;
}

}  // extern "C"
