// Context state of libemfdtd.so (host side). Not part of the public ABI.
#pragma once

#include "../../include/emfdtd.h"
#include "internal.h"

#include <cstdint>
#include <string>
#include <vector>

namespace emfd {

// Thread-local error reporting for the C ABI.
int fail(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define EMFD_HIP(call)                                      \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return ::emfd::hip_fail(e_, #call); \
    } while (0)

// Device buffer with capacity tracking.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);  // grows (contents not kept); returns EMFD_OK or an error code
    void release();
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
};

struct PortsState {
    std::vector<emfd_port_setup> ports;
    float averaging = 4.0f;
    float wavelength = 16.0f;
    bool has_sweep = false;
    emfd_sweep_setup sweep{};
    float cell_size = 1.0e-3f;
    int64_t record_steps = 0;
    int64_t capacity_steps = 0;
    int64_t recorded = 0;
    double start_time = 0.0;
    DevBuf d_ports, d_phasors, d_series;
    std::vector<double> dft_wavelengths;  // FP64 running DFT of V, I at these wavelengths (opt-in)
    DevBuf d_dft;                         // [port][frequency][V re, V im, I re, I im]
    bool enabled() const { return !ports.empty(); }
    bool recording() const { return has_sweep || record_steps > 0; }
};

struct LinesState {
    int n = 0;                 // quantities
    float wavelength = 16.0f;  // drive wavelength as given (the rotation uses max(wavelength, 2))
    float averaging = 4.0f;
    double start_time = 0.0;
    DevBuf d_quant, d_terms, d_weights, d_phasors, d_now;
    bool enabled() const { return n > 0; }
};

struct NtffState {
    bool enabled = false;
    emfd_far_field_setup setup{};
    NtffBox box{};
    float map_normal[3] = {0, 0, 1};
    std::vector<float> dirs;   // 4 per direction
    std::vector<int> map_index;
    float frame[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    double start_time = 0.0;
    DevBuf d_phasors;  // SoA: 24 planes of n_points floats
    DevBuf d_side;     // x-face side columns (SideCols layout)
    DevBuf d_samples;  // kNtffRun sample slots of 6 planes of n_points floats (deferred accumulation)
    int pending = 0;   // sampled steps of the current run not yet applied to the phasors
    std::vector<float> wavelengths;  // accumulated wavelengths; [0] is setup.wavelength
    int n_freq() const { return wavelengths.empty() ? 1 : static_cast<int>(wavelengths.size()); }
    float wavelength(int f) const { return wavelengths.empty() ? setup.wavelength : wavelengths[static_cast<size_t>(f)]; }
    SideCols side{};   // geometry of d_side (buf set when the Yee kernel maintains it)
    DevBuf d_dirs, d_partials, d_out, d_power;
    // radiation integrals: operand geometry of the current box and its FP16 operand buffers
    FfGeom ff{};
    size_t ff_halves = 0;
    bool ff_zeroed = false;
    DevBuf d_ahi, d_alo, d_absmax;
    // FP64 running DFT of the Huygens points at every accumulated wavelength (opt-in):
    // per wavelength 12 planes of n_points doubles; d_dft_ph: FP32 planes for the pattern
    bool dft = false;
    DevBuf d_dft, d_dft_ph;
};

}  // namespace emfd

struct emfd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool configured = false;
    emfd_grid_config cfg{};
    size_t n = 0;  // cells

    // fields: two sets (ping-pong for the fused kernel); cur = the one holding the state
    emfd::DevBuf field_buf[2];
    int cur = 0;
    emfd::DevBuf d_src_descs;  // kept sources (FP64 stepper evaluates their signals on the device)
    int storage = emfd::kStorageF32;  // field storage (EMFD_STORAGE_*)
    bool side_active = false;         // the current step's Yee kernel writes the far-field side columns
    bool side_e_valid = true;         // ... including the E components
    int scale_log2[6] = {0, 0, 0, 0, 0, 0};
    size_t field_bytes() const {
        return n * 6 * (storage == emfd::kStorageF32 ? 4 : storage == emfd::kStorageF64 ? 8 : 2);
    }
    emfd::DevBuf sigma_cell, eps_cell;  // EM_Lab's alpha channels (probes)
    emfd::DevBuf mat_id, table, prof, pml_tab;
    emfd::DevBuf class_masks;  // LatticeArgs::wire_mask then ::class_mask, built with the class table
    bool masks_ok = false;
    int id_bytes = 0;
    int classes = 0;
    int has_wire = 0;
    int pml_materials = 1;  // LatticeArgs::pml_materials
    // texels the one-pass kernel (EMFD_KERNEL_MARCH) recomputes after its vacuum update, sorted
    // lattice indices: fix_b thin-wire classes (B), fix_e classes != 0 and the E texels whose curl
    // reads a thin-wire texel's B (texel itself, +x, +y, +z)
    emfd::DevBuf fix_b, fix_e;
    int n_fix_b = 0, n_fix_e = 0;
    // the march's own material coefficients (emfd::MarchMat, automatic z-chunks): per texel local
    // class bytes, per tile (first entry, count), entries; and what the E recomputation still
    // takes with them (fix_e_rest: thin-wire neighbours, materials in the absorber, overflow)
    emfd::DevBuf march_lid, march_tiles, march_coef, fix_e_rest;
    int n_fix_e_rest = 0;
    bool march_tables = false;
    std::vector<uint32_t> march_classes;  // class of each entry (host copy, until the table exists)
    emfd::MarchMat march_mat() const {
        emfd::MarchMat m;
        if (march_tables) {
            m.lid = march_lid.as<uint8_t>();
            m.tiles = march_tiles.as<int2>();
            m.coef = march_coef.as<float>();
        }
        return m;
    }
    // absorber: EMFD_ABSORBER_GRADED (the reference's) or EMFD_ABSORBER_CPML, with the latter's
    // coefficient tables and auxiliary fields
    int absorber = 0;
    float cpml_sigma_max = 0.0f, cpml_grading = 0.0f, cpml_alpha_max = 0.0f;
    emfd::DevBuf d_cpml_coef, d_psi;
    emfd::DevBuf d_render;  // staging of emfd_render_volume's host-bound image (kept between frames)
    int prof_stride = 0;

    // sources
    std::vector<emfd_source_desc> sources;  // kept sources, in order
    emfd::DevBuf d_packed, d_cells, d_lists;
    int n_cells = 0;
    // per source texel records (launch_build_source_records), rebuilt before a step after the
    // sources or the materials change
    emfd::DevBuf d_src_recs, d_src_pairs;
    size_t sources_lists_len = 0;  // entries of d_lists (one SourcePair each)
    bool src_recs_dirty = true;

    // per-call step constants: pinned host staging (two buffers, each released by an event once
    // its copy has run) and the device copy the kernels read
    emfd::DevBuf d_consts;
    // FP64 rotations of the running DFTs for one step call: [step][frequency][4], ports' first
    std::vector<double> h_consts64;
    emfd::DevBuf d_consts64;
    float *h_consts[2] = {nullptr, nullptr};
    size_t h_consts_cap[2] = {0, 0};
    hipEvent_t h_consts_ev[2] = {nullptr, nullptr};
    int h_consts_next = 0;

    uint64_t step_count = 0;
    int kernel_variant = EMFD_KERNEL_AUTO;
    int tile_cfg = 0;

    // kernel timing: the Yee launch of every timing_every-th step is bracketed (0 = off)
    int timing_every = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    double timing_ms = 0.0;
    int64_t timing_launches = 0;
    std::vector<float> timing_samples;  // each bracketed launch's milliseconds since switched on

    emfd::PortsState ports;
    emfd::LinesState lines;
    emfd::NtffState ntff;

    emfd::DevBuf scratch_partials, scratch_misc;

    emfd::FieldPtrs fields(int which) const;
    emfd::LatticeArgs lattice() const;
    double time() const { return static_cast<double>(step_count) * static_cast<double>(cfg.courant); }
};

namespace emfd {
// Per-step monitor hooks used by emfd_grid_step (monitors.cpp).
int monitors_consts_floats(emfd_ctx *ctx);
int monitors_fill_consts(emfd_ctx *ctx, float *consts_step, uint64_t step_after);
int monitors_after_step(emfd_ctx *ctx, const float *consts_step, int stride, int k, int count);
int monitors_dft_freqs(const emfd_ctx *ctx);
int monitors_upload_dft(emfd_ctx *ctx, int count);
}  // namespace emfd
