/*
 * emfdtd.h — C ABI of the MI355X FDTD engine (libemfdtd.so).
 *
 * This is the drop-in boundary for EM_Lab's solver path. EM_Lab's host classes bind GL texture
 * names directly; here every class of that path becomes an opaque context plus plain-pointer
 * entry points, so a C++ (or any FFI) caller can keep EM_Lab's call sequence:
 *
 *   SimulationGrid::configure / uploadMaterials / setSources / resetFields / step
 *       -> emfd_grid_configure / emfd_grid_upload_materials / emfd_grid_set_sources /
 *          emfd_grid_reset_fields / emfd_grid_step
 *   PortMonitor, FarFieldMonitor, FieldAnalyzer -> emfd_ports_*, emfd_ntff_*, emfd_flux_*, emfd_probe_*
 *
 * Conventions
 *   - Every function returns EMFD_OK (0) or a positive error code; emfd_last_error() gives a
 *     thread-local message for the last failure on the calling thread. No exception crosses
 *     the ABI.
 *   - A context is bound to one HIP device and one HIP stream (default priority) that it
 *     creates. It is not thread-safe, like the GL context it replaces.
 *   - Host arrays use EM_Lab's lattice index (z*ny + y)*nx + x (x fastest,
 *     src/SimulationGrid.hpp:55-57). Field downloads offer EM_Lab's texel layout (RGBA per
 *     cell, alpha = cell sigma for E and cell eps_r for B) or structure-of-arrays.
 *   - Silent clamping follows EM_Lab where it clamps (configure, sources, ports); sizes it
 *     rejects with std::runtime_error are rejected here with EMFD_ERR_INVALID.
 */
#ifndef EMFDTD_H
#define EMFDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMFD_OK 0
#define EMFD_ERR_INVALID 1 /* bad argument or size mismatch */
#define EMFD_ERR_HIP 2     /* HIP runtime failure (message has the HIP error string) */
#define EMFD_ERR_STATE 3   /* call out of order (e.g. step before configure) */
#define EMFD_ERR_NOMEM 4   /* device allocation failed */

#define EMFD_LAYOUT_TEXEL 0 /* float[n][4]: (x, y, z, alpha) per cell, EM_Lab's RGBA32F texel */
#define EMFD_LAYOUT_SOA 1   /* float[3][n]: component planes */

#define EMFD_KERNEL_AUTO 0     /* fastest variant that supports the grid (default): EMFD_KERNEL_MARCH for FP32
                                   lattices of at least 8e6 cells, else the vectorised two-pass kernel */
#define EMFD_KERNEL_TWO_PASS 2 /* separate update_B / update_E passes, in place */
#define EMFD_KERNEL_TWO_PASS4 5 /* two passes, four cells per thread with float4 streams and non-temporal
                                   hints (nx % 4 == 0) */
#define EMFD_KERNEL_MARCH 8     /* fused B+E in one pass (ping-pong buffers, nx % 4 == 0, planes of at most
                                   2 GB): z-march of 256-cell x segments of 7 rows with one workgroup barrier
                                   per plane, vacuum and absorber coefficients in the march, then sparse
                                   passes recompute the texels of non-vacuum classes and next to thin wires;
                                   bits 8.. of the variant = z-chunk (0 = automatic); also runs 16-bit
                                   storage (rounding B^{n+1/2} as stored), where it is not automatic */

#define EMFD_ABSORBER_GRADED 0 /* the reference's graded absorber (default; bit-exact parity) */
#define EMFD_ABSORBER_CPML 1   /* convolutional PML with auxiliary fields in the absorber slabs */

#define EMFD_STORAGE_F32 0  /* FP32 field planes (default; EM_Lab's RGBA32F) */
#define EMFD_STORAGE_F16 1  /* binary16 planes of value / 2^k per component, FP32 updates */
#define EMFD_STORAGE_BF16 2 /* bfloat16 planes of value / 2^k per component, FP32 updates */
#define EMFD_STORAGE_F64 3  /* FP64 planes and FP64 updates: the accuracy reference */

typedef struct emfd_ctx emfd_ctx;

/* GridConfig (src/SimulationGrid.hpp:25-31) */
typedef struct {
    int size[3];
    float courant;
    int pml_thickness;
    float pml_sigma_max;
    float pml_grading;
} emfd_grid_config;

/* EdgeOverride (src/SimulationGrid.hpp:76-81) */
typedef struct {
    int texel[3];
    int axis;
    float sigma;
} emfd_edge_override;

/* WireCoefficient (src/SimulationGrid.hpp:86-94) */
typedef struct {
    int texel[3];
    int component;
    int term;
    float factor;
} emfd_wire_coefficient;

/* SourceDesc (src/SimulationGrid.hpp:113-150); waveform 0 = CW, 1 = Gaussian pulse */
typedef struct {
    float center[3];
    float half_extent[3];
    float softness[3];
    float polarization[3];
    float amplitude;
    float wavelength;
    float phase;
    int waveform;
    float ramp_periods;
    float repeat_periods;
    float delay_periods;
    int enabled;
} emfd_source_desc;

/* PortSetup (src/Ports.hpp:30-40) */
typedef struct {
    int texel[3];
    int axis;
    float sign;
    int edges; /* 1 = thin-wire gap edge, 4 = rod gap corner edges */
    float wavelength;
    int voltage_feed;
    float impedance_ohm;
    float drive;
} emfd_port_setup;

/* SweepSetup (src/Ports.hpp:42-47) */
typedef struct {
    float wavelength_min;
    float wavelength_max;
    int points;
} emfd_sweep_setup;

/* PortReading (src/Ports.hpp:49-58) */
typedef struct {
    int valid;
    double v_re, v_im, i_re, i_im;
    double z_re, z_im;
    double gamma_re, gamma_im;
    double vswr;
    double return_loss_db;
    double accepted_power;
    double available_power;
} emfd_port_reading;

/* SweepPoint (src/Ports.hpp:60-65) */
typedef struct {
    double wavelength;
    double z_re, z_im;
    double s11_db;
    double vswr;
} emfd_sweep_point;

/* PortSweep summary (src/Ports.hpp:67-74) */
typedef struct {
    int64_t best;
    double best_wavelength;
    int has_band;
    double band_low_wavelength;
    double band_high_wavelength;
} emfd_port_sweep;

/* FarFieldSetup (src/FarField.hpp:39-47) */
typedef struct {
    int box_min[3];
    int box_max[3];
    float wavelength;
    float averaging_periods;
    float map_axis[3];
} emfd_far_field_setup;

/* FluxSurface (src/FieldAnalyzer.hpp:32-51); kind 0 = plane, 1 = closed box */
typedef struct {
    int kind;
    int axis;
    int index;
    int rect_min[2];
    int rect_max[2];
    int box_min[3];
    int box_max[3];
} emfd_flux_surface;

/* ProbeSample (src/FieldAnalyzer.hpp:19-29) */
typedef struct {
    float position[3];
    float E[3];
    float B[3];
    float S[3];
    float sigma;
    float epsilon_r;
} emfd_probe_sample;

/* ---------------------------------------------------------------------------------------
 * Context
 * --------------------------------------------------------------------------------------- */
/* Replaces SimulationGrid::SimulationGrid(shaderDir) (src/SimulationGrid.cpp:135-147). */
int emfd_create(int device, emfd_ctx **out);
void emfd_destroy(emfd_ctx *ctx);
const char *emfd_last_error(void);
/* Library build string (kernel variants, target). */
const char *emfd_version(void);

/* ---------------------------------------------------------------------------------------
 * SimulationGrid (src/SimulationGrid.hpp:155-216)
 * --------------------------------------------------------------------------------------- */
/* SimulationGrid::configure (src/SimulationGrid.cpp:149-199): same clamps; vacuum materials,
 * zero fields, step count 0. `applied` (may be NULL) receives the clamped configuration. */
int emfd_grid_configure(emfd_ctx *ctx, const emfd_grid_config *requested, emfd_grid_config *applied);
/* SimulationGrid::uploadMaterials (src/SimulationGrid.cpp:209-295): sigma/eps_r per cell
 * (n_cells must equal the grid's cell count), edge averaging, sub-cell overrides, binary16
 * quantisation of edge sigma (clamped to 6e4), edge eps_r and thin-wire factors. Clears the
 * fields and rewinds time, as the reference does. */
int emfd_grid_upload_materials(emfd_ctx *ctx, const float *sigma, const float *eps_r, int64_t n_cells,
                               const emfd_edge_override *edges, int64_t n_edges,
                               const emfd_wire_coefficient *wires, int64_t n_wires);
/* SimulationGrid::setSources (src/SimulationGrid.cpp:297-357): disabled, zero-polarisation and
 * beyond-256 sources are skipped; *n_active (may be NULL) receives the kept count. */
int emfd_grid_set_sources(emfd_ctx *ctx, const emfd_source_desc *sources, int n, int *n_active);
/* SimulationGrid::resetFields (src/SimulationGrid.cpp:359-366). */
int emfd_grid_reset_fields(emfd_ctx *ctx);
/* SimulationGrid::step (src/SimulationGrid.cpp:368-433): count <= 0 is a no-op. Each step runs
 * B, E, sources, then every configured monitor's per-step accumulation on the device; there is
 * no host callback per step. Asynchronous on the context stream. */
int emfd_grid_step(emfd_ctx *ctx, int count);
/* Wait for all queued work of the context. */
int emfd_grid_synchronize(emfd_ctx *ctx);
/* time() = step_count * dt, stepCount(), dt() (src/SimulationGrid.hpp:185-187). Any pointer may be NULL. */
int emfd_grid_state(emfd_ctx *ctx, double *time, uint64_t *step_count, float *dt);
/* The configuration in force (after clamps). */
int emfd_grid_get_config(emfd_ctx *ctx, emfd_grid_config *out);
/* Field readback (replaces eFieldTexture()/bFieldTexture() reads). Either pointer may be NULL.
 * Sizes: 4*n floats (TEXEL) or 3*n floats (SOA). Synchronises the stream. */
int emfd_grid_download_fields(emfd_ctx *ctx, float *E, float *B, int layout);
/* Overwrite the field state (restart or test set-up). The alpha channel of a TEXEL upload is
 * ignored (materials are unchanged). */
int emfd_grid_upload_fields(emfd_ctx *ctx, const float *E, const float *B, int layout);
/* Device bytes held by the grid (fields, coefficients, source lists). */
int64_t emfd_grid_memory_bytes(emfd_ctx *ctx);
/* Choose the Yee kernel variant: EMFD_KERNEL_* in the low byte; bits 8.. carry a variant
 * parameter (MARCH: z-chunk; 16-bit storage: cells per thread), 0 = automatic. Other values:
 * EMFD_ERR_INVALID. */
int emfd_grid_set_kernel(emfd_ctx *ctx, int variant);
/* Absorber of the lattice's outer pml_thickness cells (no counterpart in the reference, whose
 * absorber is the graded conductivity of src/SimulationGrid.cpp:149-207 / shaders/common.glsl:23-54,
 * EMFD_ABSORBER_GRADED, the default after emfd_grid_configure). EMFD_ABSORBER_CPML switches the
 * graded conductivity off and adds the convolutional PML (Roden-Gedney recursive convolution,
 * complex-frequency-shifted stretch, kappa = 1) in the same slabs: sigma = sigma_max d^grading,
 * alpha = alpha_max (1 - d) over the normalised depth d; sigma_max <= 0 -> 0.8 (grading + 1),
 * grading <= 0 -> 3, alpha_max <= 0 -> 0. It keeps 4 auxiliary planes per slab cell and needs
 * FP32 storage; its corrections follow either Yee form (EMFD_KERNEL_AUTO: the one-pass march from
 * 8e6 cells when every dimension exceeds 2 pml + 2 cells, else the two passes). Clears the fields
 * and rewinds time. */
int emfd_grid_set_absorber(emfd_ctx *ctx, int kind, float sigma_max, float grading, float alpha_max);
/* Field storage (no counterpart in the reference, which keeps RGBA32F textures): EMFD_STORAGE_*
 * with scale_log2[6] = k per component (E x y z, B x y z; NULL = all 0), the stored value being
 * field / 2^k rounded to nearest even. Every update still computes in FP32; only the planes are
 * 16-bit, which halves the field traffic. 16-bit storage needs nx % 8 == 0 and runs the
 * two-pass kernel under EMFD_KERNEL_AUTO (bits 8..11 of the variant = cells per thread, 8 or 4,
 * 0 = 8), or the one-pass EMFD_KERNEL_MARCH when chosen explicitly; both give the same bits.
 * EMFD_STORAGE_F64 runs every update and source in double
 * precision (the reference for accuracy studies); monitors read its fields as float. Clears the fields and rewinds time, like
 * emfd_grid_upload_materials. Downloads and uploads convert to and from FP32 values. */
int emfd_grid_set_storage(emfd_ctx *ctx, int storage, const int *scale_log2);
/* E and B as FP64 structure-of-arrays planes (3 x n doubles each; either may be NULL); exact
 * for EMFD_STORAGE_F64, converted otherwise. */
int emfd_grid_download_fields_f64(emfd_ctx *ctx, double *E, double *B);
/* Number of distinct coefficient classes after the last material upload (0 = uniform vacuum). */
int emfd_grid_material_classes(emfd_ctx *ctx, int *classes);
/* Live timing of the Yee kernel: with enable = k > 0, the Yee launch of every later step whose
 * step count is a multiple of k (k = 1: every step) is bracketed by HIP events on the context
 * stream; enable = 0 stops, enable < 0 only reads. *total_ms / *launches (may be NULL) report
 * the sum and count of the bracketed launches since timing was last switched on. Each bracket
 * costs the stream a few microseconds, so long timed runs sample with k > 1. */
int emfd_grid_kernel_timing(emfd_ctx *ctx, int enable, double *total_ms, int64_t *launches);
/* The bracketed launches of emfd_grid_kernel_timing one by one (milliseconds, in step order,
 * since timing was last switched on; at most 2^20 kept): the first min(cap, *n) are written to
 * out, *n receives how many there are. No counterpart in the reference (its timing line is an
 * exponential mean of GPU timer queries, src/Main.cpp:1151); used for medians and spreads. */
int emfd_grid_kernel_timing_samples(emfd_ctx *ctx, float *out, int cap, int *n);

/* CoaxMonitor's device side (src/Coax.cpp:95-198, shaders/line_accumulate.comp:30-48): n
 * quantities of 4 ints (first term, term count, kind 0 = E line integral / 1 = B loop integral,
 * unused), n_terms terms of 4 ints (x, y, z, component 0-2 E / 3-5 B) with one weight each.
 * After every step each quantity's weighted sum (indices clamped to the grid) is kept as its
 * instantaneous value and folded into a two-stage running DFT at the drive wavelength (E
 * sampled at t, B at t - dt/2), as the ports. Phasors and values start at 0; n = 0 disables.
 * Quantities reaching outside the term list, kinds other than 0 / 1 and components outside
 * 0..5 are rejected (EMFD_ERR_INVALID). */
int emfd_lines_configure(emfd_ctx *ctx, const int *quantities, int n, const int *terms, const float *weights,
                         int n_terms, float wavelength, float averaging_periods);
/* Phasors (stage-1 re, im, stage-2 re, im per quantity; may be NULL) and instantaneous values
 * (may be NULL) of the n >= configured quantities. */
int emfd_lines_read(emfd_ctx *ctx, float *phasors, float *now, int n);
/* CoaxMonitor::geometryFactor (src/Coax.cpp:42-83): the square coax's lattice capacitance per
 * unit length in units of eps (Z0 = eta0 / (sqrt(eps_r) G)), from an SOR solve of the
 * cross-section; host-only. */
double emfd_coax_geometry_factor(int inner, int outer);

/* ---------------------------------------------------------------------------------------
 * Radar back-projection (RadarMonitor::computeImage's imaging loop, src/Radar.cpp:622-658)
 * --------------------------------------------------------------------------------------- */
typedef struct emfd_radar_image_setup {
    float tx[3];              /* transmit phase centre (cells) */
    float axes[9];            /* azimuth axis, elevation axis, boresight (src/Radar.hpp:186) */
    float interior_min[3];    /* simulated volume inside the absorbing layer */
    float interior_max[3];
    float wavelength;         /* carrier wavelength (cells) */
    double dt;                /* solver time step */
    double sample_dt;         /* spacing of the decimated echo samples (D dt) */
    float r0, dr;             /* range bins */
    int nr;
    float az0, daz;           /* azimuth bins (degrees) */
    int naz;
    float el0, del;           /* elevation bins (degrees) */
    int nel;
} emfd_radar_image_setup;
/* The image power [nr][nel][naz] (double; -1 where the point lies outside the interior) from the
 * compressed, decimated echoes z ([m][n_el] complex float, interleaved re, im) of n_el receive
 * elements at element_pos ([n_el][3]) with receive weights ([n_el]), and the transmit delay of
 * n_tx transmitting elements at tx_pos ([n_tx][3]) with amplitudes tx_amp ([n_tx]): per point
 * P = tx + r dir(az, el), t_i = txDelay(P) + |P - element_i| (src/Radar.cpp:86-104),
 * power = |sum_i w_i z_i(t_i) e^{j w t_i}|^2 with z linearly interpolated. Runs on the
 * context's device and stream (the grid need not be configured). */
int emfd_radar_backproject(emfd_ctx *ctx, const emfd_radar_image_setup *setup, const float *z, int64_t m, int n_el,
                           const float *element_pos, const double *weights, const float *tx_pos, const double *tx_amp,
                           int n_tx, double *power);

/* ---------------------------------------------------------------------------------------
 * PortMonitor (src/Ports.hpp:84-140)
 * --------------------------------------------------------------------------------------- */
/* PortMonitor::configure (src/Ports.cpp:183-215): at most 256 ports; sweep may be NULL;
 * record_steps > 0 records the (V, I) series without a sweep. */
int emfd_ports_configure(emfd_ctx *ctx, const emfd_port_setup *ports, int n, float averaging_periods,
                         const emfd_sweep_setup *sweep, float cell_size, int64_t record_steps);
int emfd_ports_disable(emfd_ctx *ctx);
/* Raw EMA phasors: 8 floats per port (stage1 Vre Vim Ire Iim, stage2 ...). */
int emfd_ports_read_phasors(emfd_ctx *ctx, float *out, int n);
/* Running DFTs of the port gap V and I in FP64 (no counterpart in the reference, which keeps the
 * FP32 EMA and transforms recorded series on the host): from the next step on, every step adds
 * V_n e^{-i w n dt} and I_n e^{-i w (n - 1/2) dt} at each of n wavelengths (0 to 4096, > 0; n = 0
 * switches them off), the rotation computed in FP64. Clears the sums; emfd_ports_configure
 * switches them off. EMFD_ERR_STATE without configured ports. */
int emfd_ports_set_dft(emfd_ctx *ctx, const double *wavelengths, int n);
/* The sums: 4 doubles (V re, V im, I re, I im) per [port][wavelength]. */
int emfd_ports_read_dft(emfd_ctx *ctx, double *out, int64_t n_values);
/* PortMonitor::evaluate + computeReadings (src/Ports.cpp:145-185): one reading per port. */
int emfd_ports_evaluate(emfd_ctx *ctx, emfd_port_reading *out, int n);
/* Recorded series steps so far (PortMonitor::recordedSteps). */
int64_t emfd_ports_recorded_steps(emfd_ctx *ctx);
/* PortMonitor::readSeries (src/Ports.cpp:278-283): 2 floats (V, I) per [step][port];
 * cap_floats bounds the write. */
int emfd_ports_read_series(emfd_ctx *ctx, float *out, int64_t cap_floats);
/* PortMonitor::evaluateSweep (src/Ports.cpp:187-278). points: n_ports * sweep points entries,
 * port-major; summary: n_ports entries. Any output pointer may be NULL. Returns EMFD_ERR_STATE
 * when no sweep is configured or fewer than 16 steps are recorded (the reference returns). */
int emfd_ports_evaluate_sweep(emfd_ctx *ctx, double dt, emfd_sweep_point *points, emfd_port_sweep *summary,
                              double *residual, int64_t *samples, int *full);

/* ---------------------------------------------------------------------------------------
 * FarFieldMonitor (src/FarField.hpp:75-133)
 * --------------------------------------------------------------------------------------- */
/* FarFieldMonitor::configure (src/FarField.cpp:259-307): box clamped to [L+1, N-L-2];
 * EMFD_ERR_INVALID when it becomes degenerate. `applied` and `n_points` may be NULL. */
int emfd_ntff_configure(emfd_ctx *ctx, const emfd_far_field_setup *setup, emfd_far_field_setup *applied,
                        int *n_points);
int emfd_ntff_disable(emfd_ctx *ctx);
/* FarFieldMonitor::resetAccumulation (src/FarField.cpp:321-325). */
int emfd_ntff_reset(emfd_ctx *ctx);
/* Phasors in EM_Lab's PhasorPair layout: 32 floats per point (stage1 eRe eIm hRe hIm, stage2
 * ..., vec4 each with w = 0). */
int emfd_ntff_read_phasors(emfd_ctx *ctx, float *out, int64_t n_points);
/* EM_Lab's direction set (src/FarField.cpp:201-257): 10582 directions as 4 floats each.
 * map_index (may be NULL): 2*81*81 ints. frame (may be NULL): normal, u, v (9 floats). */
int emfd_ntff_directions(emfd_ctx *ctx, float *dirs, int cap, int *n_dirs, int *map_index, float *frame);
/* Radiation integrals (ntff_pattern + ntff_reduce, shaders/ntff_pattern.comp:45-110,
 * shaders/ntff_reduce.comp:19-39) for arbitrary directions (4 floats each, xyz used):
 * f2[d] = |N_perp - r x L|^2 at the monitor's wavelength; *power = radiated power P. */
int emfd_ntff_eval(emfd_ctx *ctx, const float *dirs, int n_dirs, float *f2, double *power);
/* Several far-field wavelengths at once (no counterpart in the reference, whose far_field{} has one
 * wavelength): the monitor accumulates the two-stage phasors of every wavelength in the list (1 to
 * 8, each clamped to >= 2 as configure does) from the same field samples; wavelengths[0] becomes
 * the monitor's wavelength, used by emfd_ntff_evaluate. Clears the accumulation. Every wavelength's
 * phasors are those a monitor configured with that wavelength alone would hold. */
int emfd_ntff_set_wavelengths(emfd_ctx *ctx, const float *wavelengths, int n);
/* emfd_ntff_eval / emfd_ntff_read_phasors for wavelength `freq` of the list. */
int emfd_ntff_eval_at(emfd_ctx *ctx, int freq, const float *dirs, int n_dirs, float *f2, double *power);
int emfd_ntff_read_phasors_at(emfd_ctx *ctx, int freq, float *out, int64_t n_points);
/* Running DFTs of the Huygens-node E and H in FP64 at every far-field wavelength, beside the EMA
 * phasors (no counterpart in the reference): enable != 0 clears and starts them, 0 stops them.
 * E_n is taken at n dt, H_n at (n - 1/2) dt, the rotation stepped in FP64. emfd_ntff_configure
 * switches them off; emfd_ntff_set_wavelengths restarts them. */
int emfd_ntff_set_dft(emfd_ctx *ctx, int enable);
/* The sums of wavelength `freq`: 12 doubles per point (E re xyz, E im xyz, H re xyz, H im xyz). */
int emfd_ntff_read_dft(emfd_ctx *ctx, int freq, double *out, int64_t n_points);
/* emfd_ntff_eval_at on the running DFTs (rounded to FP32) instead of the EMA phasors. *power is
 * then the frequency-domain flux 1/2 Re of (E x H*) through the Huygens box, up to the same
 * constant as the EMA's. */
int emfd_ntff_eval_dft(emfd_ctx *ctx, int freq, const float *dirs, int n_dirs, float *f2, double *power);
/* FarFieldMonitor::evaluate + finishEvaluation (src/FarField.cpp:278-335) on the reference
 * direction set. directivity_dbi: n_dirs floats (may be NULL). cuts: 3 x 6 floats per cut
 * (peak dBi, peak angle, beamwidth, front/back, unused, unused) (may be NULL). Returns
 * *valid = 0 when P <= 1e-12 or non-finite. */
int emfd_ntff_evaluate(emfd_ctx *ctx, int *valid, double *radiated_power, float *max_directivity_dbi,
                       float *max_direction, float *directivity_dbi, float *cuts);

/* ---------------------------------------------------------------------------------------
 * FieldAnalyzer (src/FieldAnalyzer.hpp:73-130)
 * --------------------------------------------------------------------------------------- */
/* FieldAnalyzer::measureFlux (src/FieldAnalyzer.cpp:317-375) for n surfaces. */
int emfd_flux_measure(emfd_ctx *ctx, const emfd_flux_surface *surfaces, int n, double *out);
/* emfd_flux_measure without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under stream capture, tight serving loops, harnesses
 * that forbid hipMalloc). The plane list, row partials and sums live in a fixed device array per
 * GPU (1 MiB + 24 KiB, set up by emfd_create), shared by every context on that GPU: calls that
 * overlap on one GPU race. Limits: n <= 64 surfaces and at most 262144 plane rows in all (a closed
 * box has the rows of its six faces), else EMFD_ERR_INVALID. Same results as emfd_flux_measure,
 * bit for bit. */
int emfd_flux_measure_noalloc(emfd_ctx *ctx, const emfd_flux_surface *surfaces, int n, double *out);
/* FieldAnalyzer::probe / submit+collect (src/FieldAnalyzer.cpp:126-272,274-315):
 * positions: 3 floats each. */
int emfd_probe_measure(emfd_ctx *ctx, const float *positions, int n, emfd_probe_sample *out);

/* ---------------------------------------------------------------------------------------
 * Volume view (src/VolumeRenderer.hpp:21-39, src/VolumeRenderer.cpp:70-121,
 * shaders/raymarch.frag): EM_Lab's viewer binds the grid's E and B textures (E with sigma, B with
 * eps_r in alpha, src/SimulationGrid.cpp:167-173,209-222) to its raymarch shader. Here the grid's
 * device planes are handed out as they are (no copy), and the same raymarch runs as a HIP kernel
 * over them.
 * --------------------------------------------------------------------------------------- */
/* The grid state on the device, in place of eFieldTexture() / bFieldTexture()
 * (src/SimulationGrid.hpp:181-182). Planes are in lattice order (x fastest, (z ny + y) nx + x),
 * element type by storage: float (F32), _Float16 or bfloat16 of value / scale (F16, BF16), double
 * (F64). The pointers stay valid until the next step, reset, configure, material upload or
 * storage / absorber change (the fused kernel alternates two sets of planes), and reading them
 * must be ordered after the context's stream (emfd_grid_synchronize, or work on that stream). */
typedef struct {
    int size[3];
    int storage;          /* EMFD_STORAGE_* */
    void *e[3];           /* E x, y, z planes */
    void *b[3];           /* B x, y, z planes */
    float scale[6];       /* value = stored * scale (E x y z, B x y z); 1 for F32 / F64 */
    const float *sigma;   /* cell sigma (the E texture's alpha), n floats */
    const float *eps_r;   /* cell eps_r (the B texture's alpha), n floats */
    int device;
    uint64_t step_count;  /* the step the planes hold */
} emfd_field_view;
int emfd_grid_field_view(emfd_ctx *ctx, emfd_field_view *out);

#define EMFD_RENDER_FIELDS 0   /* E + B glow */
#define EMFD_RENDER_E 1
#define EMFD_RENDER_B 2
#define EMFD_RENDER_POYNTING 3 /* |E x B| */
#define EMFD_RENDER_SIGNED_E 4 /* one E component, signed palette */

/* RenderSettings (src/VolumeRenderer.hpp:21-39); emfd_render_default_settings fills the
 * reference's defaults. resolutionScale is the caller's: render at the scaled size. */
typedef struct {
    int mode;               /* EMFD_RENDER_* */
    int signed_axis;        /* component for EMFD_RENDER_SIGNED_E (clamped to 0..2) */
    float gain_e, gain_b;
    float density;          /* field opacity multiplier */
    float emission;         /* field brightness multiplier */
    float step_voxels;      /* ray step in voxels (clamped to 0.1 .. 4) */
    float material_opacity;
    float exposure;
    float clip_min[3];      /* cutaway box in normalised grid coordinates (clamped to 0 .. 1) */
    float clip_max[3];
} emfd_render_settings;
void emfd_render_default_settings(emfd_render_settings *out);

#define EMFD_IMAGE_RGBA8 0      /* 4 bytes per pixel, unorm (the framebuffer's format) */
#define EMFD_IMAGE_RGBA32F 1    /* 4 floats per pixel: the fragment shader's output */
#define EMFD_IMAGE_DEVICE 0x100 /* flag: `image` is device memory; the call only queues the kernel */

/* VolumeRenderer::render (src/VolumeRenderer.cpp:70-121) with raymarch.frag: one frame of
 * width x height pixels, bottom row first (the framebuffer's order), of the current fields.
 * inv_view_proj: inverse of the camera's projection * view, 16 floats column-major (glm's
 * layout, src/Camera.hpp:36); cam_pos: the eye (src/Camera.cpp:66-68). The world box is the
 * reference's: the lattice centred on the origin, its longest axis 2 units
 * (src/VolumeRenderer.cpp:44-54). With EMFD_IMAGE_DEVICE, image is device memory on the
 * context's device and the kernel is queued on the context's stream; otherwise the call copies the
 * image to host memory and returns when it is there. */
int emfd_render_volume(emfd_ctx *ctx, const emfd_render_settings *settings, const float *inv_view_proj,
                       const float *cam_pos, int width, int height, int format, void *image);

/* ---------------------------------------------------------------------------------------
 * Batched stepper (no counterpart in the reference, which runs one grid per process,
 * src/Main.cpp:239-246): many independent scenes of uneven sizes advance together. Each scene
 * is what emfd_grid_configure + emfd_grid_upload_materials + emfd_grid_set_sources +
 * emfd_ports_configure would set up on its own context, and its fields, port phasors and
 * series are bit-identical to that single-grid run (same per-texel arithmetic). The Yee
 * passes run as one launch over a flat list of (scene, 1024..4096-cell) tiles; a step
 * sequence can be captured in a HIP graph; a launch-shape autotuner picks the tiling.
 * --------------------------------------------------------------------------------------- */
typedef struct emfd_batch emfd_batch;

typedef struct {
    emfd_grid_config config;                /* clamped as emfd_grid_configure; nx % 4 == 0 after clamping */
    const float *sigma;                     /* cell sigma / eps_r (n cells each), or both NULL for vacuum */
    const float *eps_r;
    int64_t n_cells;
    const emfd_edge_override *edges;
    int64_t n_edges;
    const emfd_wire_coefficient *wires;
    int64_t n_wires;
    const emfd_source_desc *sources;
    int n_sources;
    const emfd_port_setup *ports;           /* at most EMFD_BATCH_MAX_PORTS */
    int n_ports;
    float averaging_periods;
    const emfd_sweep_setup *sweep;          /* NULL: no sweep for this scene */
} emfd_batch_scene;

#define EMFD_BATCH_MAX_PORTS 8

/* Launch shape of the batched Yee passes. */
typedef struct {
    int quads_per_thread;  /* 1, 2 or 4: a tile is 256 threads x this many 4-cell x-quads */
    int persistent_blocks; /* 0: one workgroup per tile; > 0: this many workgroups loop over the tiles */
    int xcd_slices;        /* 1: plain order; 8: the tile list is cut in 8 contiguous slices, one per XCD */
} emfd_batch_launch;

int emfd_batch_create(int device, emfd_batch **out);
int emfd_batch_destroy(emfd_batch *b);
/* Adds a scene (allocates its device buffers and uploads its materials); *index = its number.
 * `applied` (may be NULL) receives the clamped configuration. */
int emfd_batch_add_scene(emfd_batch *b, const emfd_batch_scene *scene, int *index, emfd_grid_config *applied);
/* Builds the tile list and the per-step constant table for max_steps steps (source signals and
 * port rotations, FP64 on the host cast to float, as emfd_grid_step) and the port series
 * buffers. Scenes cannot be added afterwards. */
int emfd_batch_finalize(emfd_batch *b, int64_t max_steps);
int emfd_batch_info(emfd_batch *b, int *n_scenes, int64_t *total_cells, int64_t *n_tiles, size_t *device_bytes);
int emfd_batch_set_launch(emfd_batch *b, const emfd_batch_launch *launch);
int emfd_batch_get_launch(emfd_batch *b, emfd_batch_launch *launch);
/* Times each candidate launch shape over `iterations` Yee steps and keeps the fastest. Only
 * before the first step: the fields are zero then, and the Yee passes map zero to zero, so the
 * state is unchanged. ms_per_step: 3 doubles per candidate tried (may be NULL, cap entries). */
int emfd_batch_autotune(emfd_batch *b, int iterations, emfd_batch_launch *chosen, double *ms_per_step, int cap,
                        int *n_tried);
/* steps_per_graph > 0: emfd_batch_step replays a HIP graph of that many steps (captured on
 * first use) for whole chunks and launches the remainder directly; 0: direct launches only. */
int emfd_batch_set_graph(emfd_batch *b, int steps_per_graph);
/* Advances every scene by `count` steps (B, E, sources, ++step, ports: emfd_grid_step's order). */
int emfd_batch_step(emfd_batch *b, int count);
int emfd_batch_synchronize(emfd_batch *b);
int emfd_batch_state(emfd_batch *b, uint64_t *step_count);
/* Zero all fields, port phasors and series; rewind to step 0. */
int emfd_batch_reset(emfd_batch *b);
/* Scene fields as EMFD_LAYOUT_SOA planes (3 x n floats each; either pointer may be NULL). */
int emfd_batch_download_fields(emfd_batch *b, int scene, float *E, float *B);
int emfd_batch_upload_fields(emfd_batch *b, int scene, const float *E, const float *B);
int emfd_batch_read_phasors(emfd_batch *b, int scene, float *out, int n);
/* (V, I) per recorded step and port: out[2 * (k * n_ports + p) + {0, 1}]. */
int emfd_batch_read_series(emfd_batch *b, int scene, float *out, int64_t n);
/* PortMonitor::evaluateSweep for every scene with a sweep: the FP64 DFT of the recorded series
 * runs on the GPU (one workgroup per scene, one lane per frequency, the same rotation recurrence
 * renormalised every 1024 samples); Z, S11, VSWR and the band summary follow on the host as in
 * emfd_ports_evaluate_sweep. Results are kept for emfd_batch_sweep_result. */
int emfd_batch_evaluate_sweeps(emfd_batch *b);
/* points: n_ports x sweep points of the scene (may be NULL); summary: n_ports entries (may be
 * NULL); residual: ring-down ratio of port 0 (may be NULL). */
int emfd_batch_sweep_result(emfd_batch *b, int scene, emfd_sweep_point *points, int cap_points,
                            emfd_port_sweep *summary, double *residual);
/* Yee launch timing as emfd_grid_kernel_timing (both passes of a step form one launch). */
int emfd_batch_kernel_timing(emfd_batch *b, int enable, double *total_ms, int64_t *launches);

/* ---------------------------------------------------------------------------------------
 * MLP surrogate (no counterpart in the reference): a fully connected network with GELU hidden
 * layers, BF16 weights and activations, FP32 accumulation on the matrix cores, FP32 master
 * parameters and Adam. Forward, backward and the optimizer are this library's own kernels.
 * Parameter layout (host side): for each layer, W [n_out][n_in] row-major, then b [n_out].
 * The loss is the mean squared error over batch x n_out outputs.
 * --------------------------------------------------------------------------------------- */
typedef struct emfd_mlp emfd_mlp;

/* dims: n_layers + 1 widths, input first (e.g. 18, 256, 256, 256, 128); max_batch: rows per
 * device pass (forward calls with more rows are processed in chunks). */
int emfd_mlp_create(int device, const int *dims, int n_dims, int max_batch, emfd_mlp **out);
int emfd_mlp_destroy(emfd_mlp *m);
int emfd_mlp_param_count(emfd_mlp *m, int64_t *n);
/* Sets the FP32 master parameters (and their BF16 copies); resets the Adam state. */
int emfd_mlp_set_params(emfd_mlp *m, const float *params, int64_t n);
int emfd_mlp_get_params(emfd_mlp *m, float *params, int64_t n);
/* y [batch][n_out] = network(x [batch][n_in]). */
int emfd_mlp_forward(emfd_mlp *m, const float *x, int64_t batch, float *y);
/* Loss and parameter gradients (host layout) of one batch (batch <= max_batch); grads may be NULL. */
int emfd_mlp_gradients(emfd_mlp *m, const float *x, const float *t, int64_t batch, float *grads, double *loss);
int emfd_mlp_set_adam(emfd_mlp *m, float beta1, float beta2, float eps, float weight_decay);
/* One Adam step on a batch (batch <= max_batch); *loss (may be NULL) = loss before the step. */
int emfd_mlp_train_step(emfd_mlp *m, const float *x, const float *t, int64_t batch, float lr, double *loss);
/* Keeps a training set on the device (x [n][n_in], t [n][n_out]). */
int emfd_mlp_set_dataset(emfd_mlp *m, const float *x, const float *t, int64_t n);
/* `epochs` passes over the device training set in seeded random minibatches of `batch` rows
 * (the last partial minibatch included) with learning rate lr_start decaying to lr_end
 * (cosine); epoch_loss (may be NULL, cap entries): mean minibatch loss per epoch. */
int emfd_mlp_train(emfd_mlp *m, int batch, int epochs, float lr_start, float lr_end, uint64_t seed,
                   double *epoch_loss, int cap);
/* Mean time of one forward pass over `batch` device-resident rows (batch <= max_batch), from
 * HIP events over `iterations` passes. */
int emfd_mlp_time_forward(emfd_mlp *m, int64_t batch, int iterations, double *avg_ms);

/* ---------------------------------------------------------------------------------------
 * Measurement helper (no counterpart in the reference: it calibrates the roofline figures)
 * --------------------------------------------------------------------------------------- */
/* Copies `bytes` (rounded down to 16) between two scratch buffers `iterations` times on the
 * context stream with a plain float4 kernel; *avg_ms = mean time per copy from HIP events. */
int emfd_copy_probe(emfd_ctx *ctx, int64_t bytes, int iterations, double *avg_ms);
/* Stream probe (no counterpart in the reference): `planes` planes of about plane_bytes each read
 * and as many written at the same index (planes = 1, 2, 3 or 6; 6 is the fused Yee pass's access
 * pattern without its stencil), mean milliseconds per pass over `iterations`. Moves
 * 2 * planes * plane_bytes (rounded down to 32 MiB per plane) per pass. */
int emfd_stream_probe(emfd_ctx *ctx, int64_t plane_bytes, int planes, int iterations, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
