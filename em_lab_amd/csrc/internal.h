// Internal declarations shared by the host side (grid.cpp, monitors.cpp) and the device side
// (kernels.hip) of libemfdtd.so. Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/emfdtd.h"

#include <cstdint>
#include <string>
#include <vector>

namespace emfd {

// Coefficient class table entry: one per distinct (edge sigma, edge eps_r, thin-wire factors)
// combination of a lattice texel. Fields are float offsets inside a kTableStride-float record.
constexpr int kTableStride = 20;
enum TableField : int {
    kSig = 0,    // 3: max(edge sigma, 0) per E component (from binary16)
    kEps = 3,    // 3: max(edge eps_r, 1)
    kCa = 6,     // 3: exp(-x) with x = sigma dt / eps   (interior, no absorber)
    kCb = 9,     // 3: (dt / eps) phi(x)
    kWire = 12,  // 6: thin-wire factors (Bx t0, Bx t1, By t0, By t1, Bz t0, Bz t1)
};

// Side copies of the lattice columns x = X - 1 and x = X for the two far-field box faces normal to
// x (X = box min / max x): the Yee passes (and the source pass) write each value there as they
// store it, so the far-field accumulation reads those faces coalesced instead of gathering one or
// two floats per cache line. Layout [face][component 0..5 (E, B)][x - (X - 1)][z - z0][y - y0].
// Layout [face][component E x y z, B x y z][column x = X - 1, X][z - z0][yp]: y in groups of
// kSideGroup rows (the one-pass march's row groups, y / 7) padded to 8 slots, yp = 8 (y / 7 -
// gy0) + y % 7, so that the march stores a group's values of one plane as one aligned 32-byte
// sector.
constexpr int kSideGroup = 7;
struct SideCols {
    float *buf;   // null: not maintained
    int x[2];     // X of face 0 and face 1
    int y0, ny;   // y range (box min y - 1 .. box max y)
    int z0, nz;   // z range
    int e_valid;  // the E components are maintained too (else only B: read E from the planes)
    int gy0;      // first row group: y0 / 7
    int nyp;      // padded slots per z: 8 x row groups
};
__host__ __device__ __forceinline__ int side_yp(const SideCols &s, int y) {
    return (y / kSideGroup - s.gy0) * 8 + y % kSideGroup;
}

// Per-grid constants every lattice kernel needs.
struct LatticeArgs {
    int nx, ny, nz;
    int pml;
    float dt;
    int prof_stride;            // >= max(nx, ny, nz)
    const float *prof;          // absorber profile [axis][half][i]: sigma_pml contribution of axis
                                // at coordinate i (half = 0) or i + 1/2 (half = 1)
    const void *mat_id;         // per-texel class (uint16 or uint32), or null for uniform vacuum
    int id_bytes;               // 2 or 4
    const float *table;         // classes x kTableStride floats
    int has_wire;               // scene has thin wires
    const float *pml_tab;       // absorber coefficients per (kx, ky, kz) depth class, kPmlStride floats
    int pml_k;                  // depth classes per axis: 2 * pml + 2
    SideCols side;              // far-field x-face columns (buf null when not maintained)
    const uint32_t *wire_mask;  // 1 bit per texel: its class has a thin-wire factor != 1 (or null);
                                // rows of mask_wpr words, bit x % 32 of word x / 32
    const uint32_t *class_mask; // same layout: the texel's class is not 0 (vacuum), or null
    int mask_wpr;
    int pml_materials;          // some texel of class != 0 lies in the absorber (0 only when known)
};

// Writes value(c, j) as component comp0 + c (c = 0..2) of the cells x0 + j (j = 0..N-1) of row
// (y, z) that are side columns: the index arithmetic of side_index() hoisted out of the cell loop,
// since every wavefront whose rows cross the box runs this (the lanes at the faces divergently).
template <int N, class F>
__device__ __forceinline__ void side_write_cells(const SideCols &s, int comp0, int x0, int y, int z, F value) {
    if (!s.buf || y < s.y0 || y >= s.y0 + s.ny || z < s.z0 || z >= s.z0 + s.nz) return;
    const long col = static_cast<long>(s.nz) * s.nyp;  // one x column of one component
    const long yz = static_cast<long>(z - s.z0) * s.nyp + side_yp(s, y);
#pragma unroll
    for (int face = 0; face < 2; ++face) {
        const int xo0 = x0 - (s.x[face] - 1);  // column offset of cell j: xo0 + j, kept for 0 or 1
        if (xo0 > 1 || xo0 + N - 1 < 0) continue;
        const long kf = static_cast<long>(face * 6 + comp0) * 2 * col + yz;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int xo = xo0 + j;
            if (xo < 0 || xo > 1) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c) s.buf[kf + (2 * c + xo) * col] = value(c, j);
        }
    }
}

// Index of (face, comp, x, y, z) in the side columns, or -1 when (x, y, z) is not a side column.
__host__ __device__ inline long side_index(const SideCols &s, int face, int comp, int x, int y, int z) {
    const int xo = x - (s.x[face] - 1);
    if (xo < 0 || xo > 1 || y < s.y0 || y >= s.y0 + s.ny || z < s.z0 || z >= s.z0 + s.nz) return -1;
    return ((static_cast<long>(face * 6 + comp) * 2 + xo) * s.nz + (z - s.z0)) * s.nyp + side_yp(s, y);
}

// Absorber coefficient table entry (vacuum edges): B: exp(-s dt), dt phi(s dt) per component;
// E: Ca, Cb per component for sigma = sigma_pml, eps_r = 1. The depth class of lattice index i
// on an axis of n cells is 0 inside, 1 + i below the layer's inner edge, 1 + L + (i - (n-1-L))
// from the upper layer's first +1/2 sample on.
constexpr int kPmlStride = 16;
enum PmlField : int { kPmlBa = 0, kPmlBb = 3, kPmlEa = 6, kPmlEb = 9 };

// Field storage of a lattice: FP32 planes, or 16-bit planes (binary16 or bfloat16) holding
// value / scale with a power-of-two scale per component (updates still run in FP32).
enum FieldStorage : int { kStorageF32 = 0, kStorageF16 = 1, kStorageBF16 = 2, kStorageF64 = 3 };

struct FieldPtrs {
    float *e[3];  // component planes; for 16-bit / FP64 storage they point at those elements
    float *b[3];
    int storage = kStorageF32;
    float scale[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};  // E x y z, B x y z: value = stored * scale
    float inv_scale[6] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f};  // 1 / scale (exact: powers of two)
};

// Device-side step constants: per step k, nsrc signals, then port rotation (4), then the
// far-field rotation (4). Stride per step = step_stride floats.
struct StepConsts {
    const float *base;
    int step_stride;
    int nsrc;
    int port_off;
    int ntff_off;
};

// Radar back-projection image (radar.hip): power [nr][nel][naz] on the device.
hipError_t launch_radar_backproject(const emfd_radar_image_setup &s, const float *z, long m, int n_el,
                                    const float *el_pos, const double *el_w, const float *tx_pos,
                                    const double *tx_amp, int n_tx, double *power, hipStream_t st);

// Volume view of the fields (render.hip): one frame of EM_Lab's emission / absorption raymarch.
struct RenderArgs {
    float inv_vp[16];  // inverse view-projection, column-major: element (row r, column c) at [4 c + r]
    float cam[3];
    float vol_min[3], vol_size[3];    // world box of the whole lattice
    float clip_min[3], clip_max[3];   // world box marched
    int n[3];                         // cells per axis
    float step_voxels, gain_e, gain_b, density, emission, material_opacity, exposure;
    int mode, signed_axis;
    int width, height;
};
// out: width x height pixels, bottom row first; rgba8 != 0: 4 bytes per pixel (unorm), else 4 floats
hipError_t launch_raymarch(const RenderArgs &r, const FieldPtrs &f, const float *sigma_cell, const float *eps_cell,
                           void *out, int rgba8, hipStream_t s);

// ---- launchers (kernels.hip) ----
hipError_t launch_build_table(const uint16_t *keys, int classes, float dt, float *table, hipStream_t s);
// Material coefficients the one-pass march applies itself (march tiles of its automatic
// z-chunk: 256 x cells, 7 rows, kMarchTileZ planes; tile = x / 256 + nseg * (y / 7 + nyt * (z / kMarchTileZ))). Each
// tile lists up to 255 of the classes of its texels outside the absorber; lid holds a texel's
// 1-based position in its tile's list (0: none, the march's vacuum / absorber coefficients).
// Entries: kMarchCoefStride floats each (Ca x3, Cb x3 of the class, 2 unused).
constexpr int kMarchTileRows = 7, kMarchTileZ = 8, kMarchTileX = 256, kMarchMaxLocal = 255;
static_assert(kMarchTileRows == kSideGroup, "side column row groups are the march's");
constexpr int kMarchCoefStride = 8;
struct MarchMat {
    const uint8_t *lid = nullptr;  // per texel, lattice order; null: no material tables
    const int2 *tiles = nullptr;   // per tile: first entry, entry count
    const float *coef = nullptr;   // entries x kMarchCoefStride floats
};

// Fused one-pass z-march (yee_march.hip): src -> dst, then the sparse recomputation of the texels
// in fix_b (B, thin-wire classes) and fix_e (E, non-vacuum classes and thin-wire neighbours).
// cfg: z-chunk (0 = automatic). With the automatic chunk and material tables (mm.lid set),
// the march applies the listed classes' coefficients and the E recomputation takes fix_e_rest
// (thin-wire neighbours, materials in the absorber, classes past a tile's 255) instead of fix_e.
// cpml (or null): the convolutional PML's tables; its B corrections follow the B recomputation,
// its E recomputation and corrections follow the E one (launch_cpml_march).
struct MarchCpml {
    const float *coef;
    float *psi;
};
hipError_t launch_yee_march(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int cfg,
                            const uint32_t *fix_b, int n_fix_b, const uint32_t *fix_e, int n_fix_e,
                            const MarchMat &mm, const uint32_t *fix_e_rest, int n_fix_e_rest,
                            const MarchCpml *cpml, hipStream_t s);
int yee_march_supported(const LatticeArgs &a);
// passes: bit 0 the Faraday (B) pass, bit 1 the Ampere (E) pass
hipError_t launch_yee_two_pass(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s, int passes = 3);
hipError_t launch_yee_two_pass4(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s, int passes = 3);
// Convolutional PML (cpml.hip): the auxiliary fields' update and correction of one phase (B after
// the Faraday pass, E after the Ampere pass) in the absorber slabs of every axis, x then y then z.
// coef: [axis][E position i / B position i + 1/2][i][b, a] over prof_stride entries; psi: per axis
// 4 components (B_v, B_w, E_v, E_w) over the slab of 2 pml + 1 planes, x fastest.
hipError_t launch_cpml(const LatticeArgs &a, const FieldPtrs &f, const float *coef, float *psi, bool e_phase,
                       hipStream_t s);
size_t cpml_psi_floats(int nx, int ny, int nz, int pml);
// The same absorber after the one-pass march (src: step n, dst: the march's output): e_phase
// false, the B corrections of the slabs; true, E recomputed from the corrected B wherever a curl
// reads one, then the E corrections. Needs every dimension above 2 pml + 2.
hipError_t launch_cpml_march(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, const float *coef,
                             float *psi, bool e_phase, hipStream_t s);
int cpml_march_supported(const LatticeArgs &a);
int yee_two_pass4_supported(const LatticeArgs &a);
hipError_t launch_yee_half(const LatticeArgs &a, const FieldPtrs &f, int cells_per_thread, hipStream_t s);
hipError_t launch_yee_f64(const LatticeArgs &a, const FieldPtrs &f, hipStream_t s);
hipError_t launch_sources_f64(const LatticeArgs &a, const FieldPtrs &f, const int *cells, int n_cells,
                              const int *lists, const float *packed, const emfd_source_desc *descs, double t,
                              hipStream_t s);
int yee_half_supported(const LatticeArgs &a);
hipError_t launch_apply_sources(const LatticeArgs &a, const FieldPtrs &f, const int *cells, int n_cells,
                                const int *lists, const float *packed, const float *signals, hipStream_t s);
// The source pass from per-texel records (SourceRec, SourcePair in sources_device.h, 64 and 16
// bytes each): launch_build_source_records fills them after sources or materials change.
constexpr size_t kSourceRecBytes = 64, kSourcePairBytes = 16;
hipError_t launch_build_source_records(const LatticeArgs &a, const int *cells, int n_cells, const int *lists,
                                       const float *packed, void *recs, void *pairs, hipStream_t s);
hipError_t launch_apply_source_records(const LatticeArgs &a, const FieldPtrs &f, const void *recs, const void *pairs,
                                       int n_cells, const float *packed, const float *signals, hipStream_t s);
hipError_t launch_fill(float *p, size_t n, float v, hipStream_t s);
hipError_t launch_copy4(const void *src, void *dst, size_t bytes, hipStream_t s);
// R = W = planes (1, 2, 3 or 6) planes of plane_f4 float4 each, read and written at one index
hipError_t launch_stream_probe(const void *src, void *dst, size_t plane_f4, int planes, hipStream_t s);
hipError_t launch_build_class_masks(const LatticeArgs &a, uint32_t *wire_mask, uint32_t *class_mask, int wpr,
                                   hipStream_t s);
hipError_t launch_build_pml_table(const LatticeArgs &a, float *tab, hipStream_t s);

struct PortDev {
    int texel[3];
    int axis;
    int edges;
    int sign;
};
hipError_t launch_port_accumulate(const LatticeArgs &a, const FieldPtrs &f, const PortDev *ports, int n_ports,
                                  const float *rot, float blend, float *phasors, float *series, int64_t record_index,
                                  hipStream_t s);

struct NtffBox {
    int bmin[3], bmax[3];
    int face_offset[7];
    int n_points;
};
// Far-field accumulation at up to kMaxNtffFreq wavelengths: phasor block f (24 planes of
// n_points floats) at phasors + 24 f n_points, rotation factors at rot + 4 f, EMA weight blend[f].
constexpr int kMaxNtffFreq = 8;
struct NtffFreqs {
    int n;
    float blend[kMaxNtffFreq];
};
// A port-monitor step (the arguments of launch_port_accumulate) carried by a far-field launch as
// extra workgroups, so that both per-step gathers take one launch; n = 0: none.
struct PortJob {
    const PortDev *ports = nullptr;
    int n = 0;
    const float *rot = nullptr;
    float blend = 0.0f;
    float *phasors = nullptr;
    float *series = nullptr;
    int64_t rec = -1;
};
hipError_t launch_ntff_accumulate(const LatticeArgs &a, const FieldPtrs &f, const NtffBox &box, const float *rot,
                                  const NtffFreqs &freqs, float *phasors, const PortJob &pj, hipStream_t s);
// transmission-line monitor: quantities and terms 4 ints each, phasors 4 floats per quantity
hipError_t launch_line_accumulate(const LatticeArgs &a, const FieldPtrs &f, const int *quantities, int n,
                                  const int *terms, const float *weights, const float *rot, float blend, float *phasors,
                                  float *now, hipStream_t s);
// deferred far-field accumulation: per-step gather into sample slot (6 planes of n_points), then
// one pass applying `steps` updates (rotation rows rot_stride floats apart) to the phasors
constexpr int kNtffRun = 16;
hipError_t launch_ntff_sample(const LatticeArgs &a, const FieldPtrs &f, const NtffBox &box, float *slot,
                              const PortJob &pj, hipStream_t s);
hipError_t launch_ntff_flush(const NtffBox &box, const float *rot, int rot_stride, int steps, const NtffFreqs &freqs,
                             const float *samples, float *phasors, hipStream_t s);
// FP64 running DFTs (dft.hip): ports [port][frequency][V re, V im, I re, I im], one launch per
// step with that step's rotations rot[frequency][cos wtE, sin wtE, cos wtH, sin wtH]; Huygens
// points per frequency 12 planes of n_points doubles (E re, E im, H re, H im, x y z each), over a
// run of deferred samples (rotation of step k, frequency q at rot[k stride + 4 q]).
hipError_t launch_port_dft(const LatticeArgs &a, const FieldPtrs &f, const PortDev *ports, int n_ports,
                           const double *rot, int n_freq, double *acc, hipStream_t s);
hipError_t launch_ntff_dft(const NtffBox &box, const double *rot, int stride, int steps, int n_freq,
                           const float *samples, double *acc, hipStream_t s);
hipError_t launch_ntff_dft_planes(const double *acc, int n_points, float *ph, hipStream_t s);

// Far-field radiation integrals as matrix products (farfield.hip). Per box face: normal axis ax
// (side 0: lower face, 1: upper), in-plane axes u (summed in the epilogue) and v (the product's
// contraction axis), U x V points from phasor index `first` (u fastest), 16-u blocks, contraction
// length kp (2 V rounded up to 16), coordinates relative to the box centre of the face plane and of
// the first u and v, and the face's offset (halves) in the FP16 operand buffers.
struct FfFace {
    int ax, side, u, v;
    int U, V, first;
    int ublocks, kp;
    float a0, u0, v0;
    size_t a_off;
};
struct FfGeom {
    FfFace f[6];
};
FfGeom ff_geometry(const NtffBox &box, size_t *a_halves);
int ff_absmax_blocks(int n_points);
// |F|^2 per direction into out_f2 and, when power_partials is not null, (n_points + 255) / 256
// FP64 partial sums of w Re(E x H*).n (the radiated power is half their sum). Scratch: absmax
// (ff_absmax_blocks floats), a_hi / a_lo (the geometry's halves each, zeroed once), face_partials
// (6 x 12 x n_dirs floats).
hipError_t launch_far_field(const NtffBox &box, const FfGeom &g, const float *phasors, const float *dirs, int n_dirs,
                            float k, float *absmax, _Float16 *a_hi, _Float16 *a_lo, float *face_partials,
                            float *out_f2, double *power_partials, hipStream_t s);

// One rectangle of a flux surface: normal axis, plane index, u / v origin and extent, and the
// index of its first row partial (one float per row of rsize[1]).
struct FluxPlane {
    int axis, index;
    int rmin[2], rsize[2];
    int first_group;
};
hipError_t launch_flux(const LatticeArgs &a, const FieldPtrs &f, const FluxPlane *planes, int n_planes,
                       int max_rows, float *partials, hipStream_t s);
// A plane's surface and outward sign (+1 / -1).
struct FluxRef {
    int surface;
    float sign;
};

// The no-allocation flux measurement (noalloc.hip): a file-scope device array per GPU holds the
// plane list, the row partials and the surface sums. Limits: 64 surfaces, 262144 rows in all.
constexpr int kFluxScratchSurfaces = 64;
constexpr int kFluxScratchRows = 1 << 18;
struct FluxScratch {
    FluxPlane *planes;  // 6 * kFluxScratchSurfaces
    FluxRef *refs;      // 6 * kFluxScratchSurfaces
    double *sums;       // kFluxScratchSurfaces
    float *partials;    // kFluxScratchRows
};
// Looks up the array's device address on `device` (once per device; call with it current).
int flux_scratch_init(int device);
// The regions of `device`'s array (flux_scratch_init must have succeeded there).
int flux_scratch(int device, FluxScratch *out);
// Per surface, in plane order: sums[s] = sum over its planes of sign * (FP64 sum of the rows'
// partials), written (not accumulated) for surfaces 0 .. n_surfaces - 1.
hipError_t launch_flux_surfaces(const FluxPlane *planes, const FluxRef *refs, int n_planes, const float *partials,
                                int n_surfaces, double *sums, hipStream_t s);
hipError_t launch_probe_gather(const LatticeArgs &a, const FieldPtrs &f, const int *origins, int n_probes,
                               const float *sigma_cell, const float *eps_cell, float *blocks, hipStream_t s);


// ---- batched stepper (batch.hip / batch.cpp) ----
// One scene of a batch as the device sees it. Quads are 4-cell x runs; a quad index q maps to
// (x0, y, z) with r = q / qx, x0 = 4 (q - r qx), z = r / ny, y = r - z ny, the divisions done as
// __umulhi by ceil(2^32 / d) (exact while q * d < 2^32, which the host checks).
struct BatchScene {
    LatticeArgs lat;
    FieldPtrs f;
    uint32_t nq, qx, mq, my;
    // sources: per-texel lists as launch_apply_sources; signals at sig_offset of the step row
    const int *src_cells;
    const int *src_lists;
    const float *src_packed;
    int sig_offset;
    // ports: rotation at rot_offset of the step row; series [step][port][2] or null
    int n_ports;
    const PortDev *ports;
    float *phasors;
    float *series;
    int rot_offset;
    float blend;
};

struct BatchLaunchArgs {
    const BatchScene *scenes;
    const int2 *tiles;  // (scene, first quad)
    int n_tiles;
    int slices;         // 1 or 8
    int quads;          // quads per thread: 1, 2 or 4
    int pml_materials;  // some scene has a material texel in its absorber (LatticeArgs::pml_materials)
};
hipError_t launch_batch_yee(const BatchLaunchArgs &a, int persistent_blocks, hipStream_t s);
hipError_t launch_batch_sources(const BatchScene *scenes, const int2 *list, int n, const float *consts,
                                int64_t stride, const int64_t *d_step, int k, hipStream_t s);
hipError_t launch_batch_ports(const BatchScene *scenes, const int2 *list, int n, const float *consts,
                              int64_t stride, const int64_t *d_step, int k, hipStream_t s);
hipError_t launch_batch_advance(int64_t *d_step, int count, hipStream_t s);

// FP64 DFT of one scene's recorded port series (PortMonitor::evaluateSweep's sums)
struct BatchSweepJob {
    const float *series;    // [n][ports][2]
    int n_ports;
    int points;
    int64_t n;              // recorded samples
    const double2 *step;    // per point: e^{-i w dt}
    double2 *sums;          // [port][point][2]: sum V rot, sum I rot
    double *residual;       // late / peak of |V| of port 0
};
hipError_t launch_batch_sweep(const BatchSweepJob *jobs, int n_jobs, hipStream_t s);

// ---- MLP surrogate (mlp.hip / mlp.cpp) ----
// BF16 GEMM with an FP32-accumulating MFMA core and a fused epilogue; see mlp.hip.
enum MlpEpilogue : int {
    MLP_EPI_STORE = 0, MLP_EPI_BIAS = 1, MLP_EPI_BIAS_GELU = 2, MLP_EPI_GELU_GRAD = 3,
    MLP_EPI_BIAS_GELU_INFER = 4  // as BIAS_GELU without keeping the pre-activation
};
struct MlpGemm {
    const void *A;  // bf16
    const void *B;  // bf16
    bool a_t, b_t;
    int M, N, K, lda, ldb;
    int epi;
    const float *bias;
    const float *aux;
    float *out_f32;
    void *out_bf16;
    int ldc;
};
hipError_t launch_mlp_gemm(const MlpGemm &m, hipStream_t s);
hipError_t launch_mlp_mse(const float *y, const float *t, size_t n, float scale, float *g, void *gb,
                          double *partials, int blocks, hipStream_t s);
hipError_t launch_mlp_colsum(const float *g, int M, int N, float *db, hipStream_t s);
hipError_t launch_mlp_gather(const float *x, int k_in, int ld, const float *t, int n_t, const int *idx, int rows,
                             void *xb, float *tb, hipStream_t s);
hipError_t launch_mlp_adam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
                           float eps, float wd, float c1, float c2, const int *bf16_index, void *wb, hipStream_t s);
hipError_t launch_mlp_to_bf16(const float *p, const int *bf16_index, size_t n, void *wb, hipStream_t s);

// Fused inference pass over every layer (mlp.hip): rows of 64 stay in LDS between layers.
constexpr int kMlpMaxLayers = 8;
struct MlpFusedArgs {
    const float *x;                   // [M][n_in] FP32
    float *y;                         // [M][n_out of the last layer] FP32
    int M, n_in, layers;
    int kp[kMlpMaxLayers];            // padded input width of each layer (multiple of 16, <= 256)
    int n_out[kMlpMaxLayers];         // output width of each layer (128 or 256)
    const __bf16 *w[kMlpMaxLayers];   // BF16 weights [n_out][kp]
    const float *bias[kMlpMaxLayers]; // FP32 biases
};
int mlp_fused_supported(int layers, const int *kp, const int *n_out);
hipError_t launch_mlp_forward_fused(const MlpFusedArgs &a, hipStream_t s);

}  // namespace emfd
