// Batched stepper kernels: many independent scenes of uneven sizes advance in one launch per
// pass. The Yee passes walk a flat list of (scene, first quad) tiles; each tile is 256 threads x
// `quads` 4-cell x runs of one scene, and every quad goes through b_quad_pre / e_quad_pre (the
// per-texel arithmetic of the single-grid vectorised kernel), so each scene's fields are
// bit-identical to a run of its own. Sources and ports walk flat (scene, item) lists with the single-grid device bodies. The
// step index lives on the device (d_step) so that a captured HIP graph can be replayed.
#include "internal.h"
#include "port_device.h"
#include "sources_device.h"
#include "yee_quad.h"

namespace emfd {
namespace {

// The tile list and the scene descriptors are written by the host before a launch and only read
// by the kernels: reading them through the constant address space lets the compiler use scalar
// loads (the tile index is uniform across the workgroup), so a tile's descriptor arrives in SGPRs
// in one short scalar round trip; and pointers read from constant memory are taken to address
// global memory, so the field and table accesses through them are global (not flat) loads and
// stores. Read through generic pointers, each tile cost three dependent vector-memory round trips
// before its first field load.
template <class T>
using ConstPtr = const __attribute__((address_space(4))) T *;
template <class T>
__device__ __forceinline__ ConstPtr<T> as_const(const T *p) {
    return reinterpret_cast<ConstPtr<T>>(reinterpret_cast<uintptr_t>(p));
}

// Member-wise copies (a struct copy cannot bind to a constant-address-space reference). Batched
// scenes keep no far-field side columns.
__device__ __forceinline__ LatticeArgs lattice_from(ConstPtr<LatticeArgs> p) {
    LatticeArgs a;
    a.nx = p->nx;
    a.ny = p->ny;
    a.nz = p->nz;
    a.pml = p->pml;
    a.dt = p->dt;
    a.prof_stride = p->prof_stride;
    a.prof = p->prof;
    a.mat_id = p->mat_id;
    a.id_bytes = p->id_bytes;
    a.table = p->table;
    a.has_wire = p->has_wire;
    a.pml_tab = p->pml_tab;
    a.pml_k = p->pml_k;
    a.side = SideCols{};
    a.wire_mask = p->wire_mask;
    a.class_mask = p->class_mask;
    a.mask_wpr = p->mask_wpr;
    return a;
}
__device__ __forceinline__ FieldPtrs fields_from(ConstPtr<FieldPtrs> p) {
    FieldPtrs f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        f.e[c] = p->e[c];
        f.b[c] = p->b[c];
    }
    f.storage = p->storage;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        f.scale[c] = p->scale[c];
        f.inv_scale[c] = p->inv_scale[c];
    }
    return f;
}

template <bool kB, int Q, bool PM>
__global__ void __launch_bounds__(256) k_batch_yee(BatchLaunchArgs args) {
    // tile t of the list is handled by the workgroups of slice t / per; with 8 slices the
    // hardware's round-robin of workgroups over the 8 XCDs keeps each slice's tiles (and their
    // shared halo rows) in one XCD's L2
    const int slices = args.slices;
    const int slice = static_cast<int>(blockIdx.x) % slices;
    const int lb = static_cast<int>(blockIdx.x) / slices;
    const int nb = static_cast<int>(gridDim.x) / slices;
    const int per = (args.n_tiles + slices - 1) / slices;
    const int t0 = slice * per;
    const int t1 = min(t0 + per, args.n_tiles);
    const ConstPtr<int2> tiles = as_const(args.tiles);
    const ConstPtr<BatchScene> scenes = as_const(args.scenes);
    for (int t = t0 + lb; t < t1; t += nb) {
        const int scene = tiles[t].x, first = tiles[t].y;
        const ConstPtr<BatchScene> sp = scenes + scene;
        const LatticeArgs lat = lattice_from(&sp->lat);
        const FieldPtrs f = fields_from(&sp->f);
        const uint32_t nq = sp->nq, qx = sp->qx, mq = sp->mq, my = sp->my;
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const uint32_t q = static_cast<uint32_t>(first) + static_cast<uint32_t>(j * 256) + threadIdx.x;
            if (q < nq) {
                const uint32_t r = __umulhi(q, mq);
                const uint32_t x0 = 4u * (q - r * qx);
                const uint32_t z = __umulhi(r, my);
                const uint32_t y = r - z * static_cast<uint32_t>(lat.ny);
                if (kB)
                    b_quad_pre<2>(lat, f, static_cast<int>(x0), static_cast<int>(y), static_cast<int>(z));
                else
                    e_quad_pre<2, PM>(lat, f, static_cast<int>(x0), static_cast<int>(y), static_cast<int>(z));
            }
        }
    }
}

__global__ void __launch_bounds__(64) k_batch_sources(const BatchScene *scenes, const int2 *list, int n,
                                                      const float *consts, int64_t stride, const int64_t *d_step,
                                                      int k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 e = list[i];
    const BatchScene &sc = scenes[e.x];
    const float *row = consts + (d_step[0] + k) * stride;
    source_cell(sc.lat, sc.f, sc.src_cells, sc.src_lists, sc.src_packed, row + sc.sig_offset, e.y);
}

__global__ void __launch_bounds__(64) k_batch_ports(const BatchScene *scenes, const int2 *list, int n,
                                                    const float *consts, int64_t stride, const int64_t *d_step,
                                                    int k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int2 e = list[i];
    const BatchScene &sc = scenes[e.x];
    const int64_t step = d_step[0] + k;  // index of the step that just completed
    const float *row = consts + step * stride;
    float *rec = sc.series ? sc.series + 2 * (step * sc.n_ports + e.y) : nullptr;
    port_sample(sc.lat, sc.f, sc.ports[e.y], row + sc.rot_offset, sc.blend, sc.phasors + 8 * e.y, rec);
}

__global__ void k_batch_advance(int64_t *d_step, int count) {
    if (threadIdx.x == 0 && blockIdx.x == 0) d_step[0] += count;
}

// One workgroup (one wavefront) per scene, one lane per sweep point. The rotation recurrence is
// PortMonitor::evaluateSweep's (src/Ports.cpp:217-233) in FP64 with products rounded as written.
__global__ void __launch_bounds__(64) k_batch_sweep(const BatchSweepJob *jobs) {
    const BatchSweepJob jb = jobs[blockIdx.x];
    const int lane = threadIdx.x;
    const int P = jb.n_ports;
    for (int p = 0; p < P; ++p) {
        for (int f = lane; f < jb.points; f += 64) {
            const double cr = jb.step[f].x, ci = jb.step[f].y;
            double rr = 1.0, ri = 0.0, svr = 0.0, svi = 0.0, sir = 0.0, sii = 0.0;
            for (int64_t k = 0; k < jb.n; ++k) {
                const float2 vi = *reinterpret_cast<const float2 *>(jb.series + 2 * (k * P + p));
                const double v = static_cast<double>(vi.x), c = static_cast<double>(vi.y);
                svr = svr + v * rr;
                svi = svi + v * ri;
                sir = sir + c * rr;
                sii = sii + c * ri;
                const double nr = rr * cr - ri * ci;
                const double ni = rr * ci + ri * cr;
                rr = nr;
                ri = ni;
                if ((k & 1023) == 1023) {
                    const double m = hypot(rr, ri);
                    rr = rr / m;
                    ri = ri / m;
                }
            }
            double2 *o = jb.sums + 2 * (static_cast<size_t>(p) * jb.points + f);
            o[0] = make_double2(svr, svi);
            o[1] = make_double2(sir, sii);
        }
    }
    // ring-down residual of port 0 (src/Ports.cpp:206-213): max is order-independent
    __shared__ double s_peak[64], s_late[64];
    double peak = 0.0, late = 0.0;
    const int64_t tail = jb.n - max(jb.n / 20, static_cast<int64_t>(1));
    for (int64_t k = lane; k < jb.n; k += 64) {
        const double v = fabs(static_cast<double>(jb.series[2 * (k * P)]));
        peak = fmax(peak, v);
        if (k >= tail) late = fmax(late, v);
    }
    s_peak[lane] = peak;
    s_late[lane] = late;
    __syncthreads();
    if (lane == 0) {
        for (int i = 1; i < 64; ++i) {
            peak = fmax(peak, s_peak[i]);
            late = fmax(late, s_late[i]);
        }
        jb.residual[0] = peak > 0.0 ? late / peak : 1.0;
    }
}

template <bool kB, bool PM>
hipError_t launch_pass_pm(const BatchLaunchArgs &a, int grid, hipStream_t s) {
    switch (a.quads) {
        case 1: k_batch_yee<kB, 1, PM><<<grid, 256, 0, s>>>(a); break;
        case 2: k_batch_yee<kB, 2, PM><<<grid, 256, 0, s>>>(a); break;
        case 4: k_batch_yee<kB, 4, PM><<<grid, 256, 0, s>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// the B pass has no absorber-material case: one instance serves both
template <bool kB>
hipError_t launch_pass(const BatchLaunchArgs &a, int grid, hipStream_t s) {
    if (kB || !a.pml_materials) return launch_pass_pm<kB, false>(a, grid, s);
    return launch_pass_pm<kB, true>(a, grid, s);
}

}  // namespace

hipError_t launch_batch_yee(const BatchLaunchArgs &a, int persistent_blocks, hipStream_t s) {
    if (a.n_tiles <= 0) return hipSuccess;
    if (a.slices != 1 && a.slices != 8) return hipErrorInvalidValue;
    int grid;
    if (persistent_blocks > 0) {
        grid = ((persistent_blocks + a.slices - 1) / a.slices) * a.slices;
    } else {
        grid = ((a.n_tiles + a.slices - 1) / a.slices) * a.slices;
    }
    hipError_t e = launch_pass<true>(a, grid, s);
    if (e != hipSuccess) return e;
    return launch_pass<false>(a, grid, s);
}

hipError_t launch_batch_sources(const BatchScene *scenes, const int2 *list, int n, const float *consts,
                                int64_t stride, const int64_t *d_step, int k, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_batch_sources<<<(n + 63) / 64, 64, 0, s>>>(scenes, list, n, consts, stride, d_step, k);
    return hipGetLastError();
}

hipError_t launch_batch_ports(const BatchScene *scenes, const int2 *list, int n, const float *consts,
                              int64_t stride, const int64_t *d_step, int k, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_batch_ports<<<(n + 63) / 64, 64, 0, s>>>(scenes, list, n, consts, stride, d_step, k);
    return hipGetLastError();
}

hipError_t launch_batch_advance(int64_t *d_step, int count, hipStream_t s) {
    k_batch_advance<<<1, 64, 0, s>>>(d_step, count);
    return hipGetLastError();
}

hipError_t launch_batch_sweep(const BatchSweepJob *jobs, int n_jobs, hipStream_t s) {
    if (n_jobs <= 0) return hipSuccess;
    k_batch_sweep<<<n_jobs, 64, 0, s>>>(jobs);
    return hipGetLastError();
}

}  // namespace emfd
