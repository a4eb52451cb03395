/*
 * CPU oracle for the EM_Lab FDTD hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C restatement of the reference's GLSL compute
 * shaders and the host code that drives them. It exists to check the HIP implementation:
 * only tests, the smoke check and the CPU-baseline leg of the benchmark may load it. The
 * product library never calls it.
 *
 * Layout follows the reference exactly: one RGBA texel per cell, x fastest,
 * index (z*ny + y)*nx + x (src/SimulationGrid.hpp:55-57).
 *   E texel: (Ex, Ey, Ez, sigma_cell)     B texel: (Bx, By, Bz, eps_cell)
 *   edge material texel: 4 x uint32 = packHalf2x16 (sx,sy), (sz,ex), (ey,ez), 0
 *   wire texel:          4 x uint32 = packHalf2x16 (Bx t0,t1), (By t0,t1), (Bz t0,t1), 0
 */
#ifndef EMLAB_ORACLE_H
#define EMLAB_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nx, ny, nz;
    float dt;        /* Courant number (normalised units) */
    int pml;         /* absorber thickness (cells) */
    float sigma_max;
    float grading;
} or_grid;

/* Source as the reference's SourceDesc (src/SimulationGrid.hpp:113-150). */
typedef struct {
    float center[3];
    float half_extent[3];
    float softness[3];
    float polarization[3];
    float amplitude;
    float wavelength;
    float phase;
    int waveform;        /* 0 = CW, 1 = Gaussian pulse */
    float ramp_periods;
    float repeat_periods;
    float delay_periods;
    int enabled;
} or_source_desc;

typedef struct { int texel[3]; int axis; float sigma; } or_edge_override;
typedef struct { int texel[3]; int component; int term; float factor; } or_wire_coef;

/* exp used for every float exponential of the path (see oracle.c) */
float or_exp(float x);

/* half floats (round to nearest even, as glm::packHalf2x16) */
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);
uint32_t or_pack_half2(float lo, float hi);

/* SimulationGrid::configure clamps (src/SimulationGrid.cpp:149-156) */
void or_configure(int nx, int ny, int nz, float courant, int pml, float sigma_max, float grading, or_grid *out);

/* common.glsl helpers (shaders/common.glsl:23-54) */
float or_pml_profile(const or_grid *g, float depth);
float or_pml_sigma(const or_grid *g, float x, float y, float z);
int or_in_pml(const or_grid *g, int x, int y, int z);
float or_phi(float x);

/* uploadMaterials (src/SimulationGrid.cpp:209-295). E/B alpha channels are written, field rgb zeroed.
 * wire may be NULL when n_wire == 0; returns 1 when the scene has thin wires (wire texture full size). */
int or_upload_materials(const or_grid *g, const float *sigma, const float *eps, const or_edge_override *edges,
                        size_t n_edges, const or_wire_coef *wires, size_t n_wire, float *E, float *B,
                        uint32_t *edge_mat, uint32_t *wire_tex);

/* setSources packing + cell lists (src/SimulationGrid.cpp:297-357).
 * packed: 16 floats per kept source; cells: 4 ints per cell + sentinel; lists: source indices.
 * Returns the number of kept sources; *n_cells / *n_list are set. Capacity checks by caller
 * (call with NULL outputs to get the sizes). kept_index[i] = index into `src` of kept source i. */
int or_set_sources(const or_grid *g, const or_source_desc *src, int n_src, float *packed, int *kept_index,
                   int *cells, int *n_cells, int *lists, int *n_list);

/* SourceDesc::signal (src/SimulationGrid.cpp:103-130), double precision */
double or_signal(const or_source_desc *s, double t);

/* Kernels, one dispatch each. */
void or_update_B(const or_grid *g, float *B, const float *E, const uint32_t *wire_tex, int thin_wires);
void or_update_E(const or_grid *g, float *E, const float *B, const uint32_t *edge_mat);
void or_apply_sources(const or_grid *g, float *E, const uint32_t *edge_mat, const float *packed, const float *signals,
                      int signal_base, const int *cells, int n_cells, const int *lists);
void or_clear_fields(const or_grid *g, float *E, float *B);

/* SimulationGrid::step for `count` steps (src/SimulationGrid.cpp:368-433), no observers.
 * step_count is advanced. */
void or_step(const or_grid *g, float *E, float *B, const uint32_t *edge_mat, const uint32_t *wire_tex, int thin_wires,
             const or_source_desc *src, int n_src, int count, uint64_t *step_count);
/* 16-bit field storage emulation (no reference counterpart): or_step with every pass's output
 * rounded to binary16 (storage 1) or bfloat16 (storage 2) of value / 2^k, k per component
 * (E x y z, B x y z), round to nearest even. */
void or_step_storage(const or_grid *g, float *E, float *B, const uint32_t *edge_mat, const uint32_t *wire_tex,
                     int thin_wires, const or_source_desc *src, int n_src, int count, uint64_t *step_count,
                     int storage, const int *scale_log2);
void or_quantize_fields(const or_grid *g, float *E, float *B, int storage, const int scale_log2[6]);
float or_round_f16(float f);
float or_round_bf16(float f);

/* ---- monitors ---- */
typedef struct { int texel[3]; int axis; int edges; int sign; } or_port;

/* port_accumulate.comp (shaders/port_accumulate.comp:52-83). phasors: 8 floats per port.
 * series (may be NULL): 2 floats per (record_index, port). */
void or_port_accumulate(const or_grid *g, const float *E, const float *B, const or_port *ports, int n_ports,
                        const float rot[4], float blend, int record_index, float *phasors, float *series);

/* PortMonitor::accumulate host constants (src/Ports.cpp:253-266): rot and blend for the step
 * whose completion leaves the grid at step_count. */
void or_line_accumulate(const or_grid *g, const float *E, const float *B, const int *quantities, int n,
                        const int *terms, const float *weights, const float rot[4], float blend, float *phasors,
                        float *now);
void or_monitor_constants(double wavelength, float averaging, float dt, uint64_t step_count, float rot[4],
                          float *blend);

/* ntff_common.glsl ntffPoint (shaders/ntff_common.glsl:30-53) */
void or_ntff_point(const int box_min[3], const int box_max[3], const int face_offset[7], int index, int node[3],
                   float normal[3], float *weight);
/* FarFieldMonitor::configure face offsets (src/FarField.cpp:274-282); returns point count */
int or_ntff_face_offsets(const int box_min[3], const int box_max[3], int face_offset[7]);
/* ntff_accumulate.comp: phasors 32 floats per point (stage1 eRe,eIm,hRe,hIm; stage2 ...), vec4 each */
void or_ntff_accumulate(const or_grid *g, const float *E, const float *B, const int box_min[3],
                        const int box_max[3], const int face_offset[7], int n_points, const float rot[4],
                        float blend, float *phasors);
/* ntff_pattern.comp + ntff_reduce.comp in the reference's float summation order: for each
 * direction, out[2d] = |F|^2, out[2d+1] = P/2 (P only for the chunk that starts at 0; the
 * reference's out[1] carries it). dirs: 4 floats each. */
void or_ntff_eval_f32(const int box_min[3], const int box_max[3], const int face_offset[7], int n_points,
                      const float *phasors, const float *dirs, int n_dirs, float k, const float center[3],
                      float *out);
/* Same integrals in double precision, plain sum (for pinning). outF2[d], *P = radiated power. */
void or_ntff_eval_f64(const int box_min[3], const int box_max[3], const int face_offset[7], int n_points,
                      const float *phasors, const double *dirs, int n_dirs, double k, const double center[3],
                      double *outF2, double *P);
/* FarFieldMonitor::buildDirections (src/FarField.cpp:201-257): writes at most `cap` directions
 * (4 floats each) and returns the full count (10590 in IEEE float: 3 cuts x 180 plus two
 * hemispheres of 5025 u-v cells). map_index: 2*81*81 ints. */
int or_build_directions(const float axis[3], float *dirs, int cap, int *map_index, float map_frame[9]);

/* field_sampling.glsl */
void or_cpml_coefficients(const or_grid *g, double sigma_max, double grading, double alpha_max, float *coef);
void or_cpml_b(const or_grid *g, float *B, const float *E, const float *coef, float *psi);
void or_cpml_e(const or_grid *g, float *E, const float *B, const uint32_t *edge_mat, const float *coef, float *psi);
void or_step_cpml(const or_grid *g, float *E, float *B, const uint32_t *edge_mat, const uint32_t *wire_tex,
                  int thin_wires, const or_source_desc *src, int n_src, int count, uint64_t *step_count,
                  const float *coef, float *psi);
int or_cpml_psi_floats(const or_grid *g);
void or_ntff_samples(const or_grid *g, const float *E, const float *B, const int box_min[3], const int box_max[3],
                     const int face_offsets[7], int n_points, float *out);
void or_e_at_node(const or_grid *g, const float *E, int x, int y, int z, float out[3]);
void or_b_at_node(const or_grid *g, const float *B, int x, int y, int z, float out[3]);

/* poynting_flux.comp + FieldAnalyzer::planeFlux (src/FieldAnalyzer.cpp:317-359) */
double or_plane_flux(const or_grid *g, const float *E, const float *B, int axis, int index, int rmin_u,
                     int rmin_v, int rmax_u, int rmax_v);
/* FieldAnalyzer::measureFlux for a closed box (src/FieldAnalyzer.cpp:361-375) */
double or_box_flux(const or_grid *g, const float *E, const float *B, const int box_min[3], const int box_max[3]);

/* probe: probe_gather + sampleBlock (src/FieldAnalyzer.cpp:19-40,274-315).
 * out: E[3], B[3], S[3], sigma, eps (11 floats) */
void or_probe(const or_grid *g, const float *E, const float *B, const float pos[3], float out[11]);

/* PortMonitor::evaluateSweep (src/Ports.cpp:187-278) for one port of a (V,I) series.
 * out per point: wavelength, Re Z, Im Z, S11 dB, VSWR (5 doubles). band[0..4] = best index,
 * best wavelength, has band, band low wavelength, band high wavelength. Returns residual. */
double or_sweep(const float *series, size_t n_steps, int n_ports, int port, double dt, double wl_min,
                double wl_max, int points, double z0, double *out, double *band);

/* Volume raymarch (render.c): shaders/raymarch.frag with the set-up of
 * src/VolumeRenderer.cpp:44-54,70-121 already applied (world box, clamps). */
typedef struct {
    float inv_vp[16];  /* column-major */
    float cam[3];
    float vol_min[3], vol_size[3], clip_min[3], clip_max[3];
    int n[3];
    float step_voxels, gain_e, gain_b, density, emission, material_opacity, exposure;
    int mode, signed_axis;
    int width, height;
} or_render_args;
float or_gradient_noise(float px, float py);
float or_texture(const float *tex, int nx, int ny, int nz, int ch, float u, float v, float w);
void or_box_hit(const float *ro, const float *rd, const float *bmin, const float *bmax, float *t_enter, float *t_exit);
void or_render_pixel(const or_render_args *r, const float *E, const float *B, int px, int py, float *rgba);
void or_render_volume(const or_render_args *r, const float *E, const float *B, float *rgba);

#ifdef __cplusplus
}
#endif
#endif
