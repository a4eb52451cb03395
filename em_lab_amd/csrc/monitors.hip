// HIP kernels of the per-step and on-demand monitors: port gap V/I with the two-stage EMA,
// near-to-far-field Huygens-box phasors, radiation integrals, Poynting flux, probe gather.
// Compiled with -ffp-contract=off, like the Yee kernels, so each expression rounds as written.
#include "internal.h"
#include "port_device.h"

namespace emfd {
namespace {

// S: the field storage format, fixed at compile time so the gathers' loads carry no format branch
template <int S>
__global__ void __launch_bounds__(64) k_port_accumulate(LatticeArgs a, FieldPtrs f, const PortDev *ports, int n_ports,
                                                        const float *rot, float blend, float *ph, float *series,
                                                        int64_t rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_ports) return;
    f.storage = S;
    port_sample(a, f, ports[i], rot, blend, ph + 8 * i,
                (rec >= 0 && series) ? series + 2 * (rec * n_ports + i) : nullptr);
}

// ---------------------------------------------------------------------------------------------
// Huygens box point (shaders/ntff_common.glsl:30-53)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void ntff_point(const NtffBox &b, int index, int node[3], float n[3], float *w) {
    int f = 0, first = 0;
#pragma unroll
    for (int i = 1; i < 6; ++i)
        if (index >= b.face_offset[i]) {
            f = i;
            first = b.face_offset[i];
        }
    const int ax = f >> 1, side = f & 1;
    const int u = ax == 0 ? 1 : 0, v = ax == 2 ? 1 : 2;
    const int nu = b.bmax[u] - b.bmin[u], nv = b.bmax[v] - b.bmin[v];
    const int local = index - first;
    const int iu = local % (nu + 1), iv = local / (nu + 1);
    // per-component selects (no runtime index into the private arrays, which would live in scratch)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        node[d] = d == ax ? (side == 1 ? b.bmax[d] : b.bmin[d]) : b.bmin[d] + (d == u ? iu : iv);
        n[d] = d == ax ? (side == 1 ? 1.0f : -1.0f) : 0.0f;
    }
    *w = ((iu == 0 || iu == nu) ? 0.5f : 1.0f) * ((iv == 0 || iv == nv) ? 0.5f : 1.0f);
}

// ntff_accumulate.comp (shaders/ntff_accumulate.comp:30-52). Phasors are stored as 24 planes of
// n_points floats: stage s (0,1) x quantity q (eRe, eIm, hRe, hIm) x component c -> plane 12 s + 3 q + c.
// Fields at Huygens point i (E, then B): the x-normal faces read the coalesced side columns when
// the Yee kernel maintains them, the other faces the planes.
__device__ __forceinline__ void ntff_fields(const LatticeArgs &a, const FieldPtrs &f, const NtffBox &box, int i,
                                            float E[3], float H[3]) {
    int node[3];
    float n[3], w;
    ntff_point(box, i, node, n, &w);
    if (a.side.buf && n[0] != 0.0f) {
        // every sample of a point on the face x = X lies in the face's columns (x = X - 1, X;
        // y, z from box min - 1 to box max), so the offsets are formed without side_index's checks,
        // which would put each load behind a branch of its own
        const SideCols &sc = a.side;
        const int face = n[0] > 0.0f ? 1 : 0;
        const float *base = sc.buf + static_cast<size_t>(face) * 12 * sc.nz * sc.nyp;
        const int x1 = sc.x[face] - 1;
        const auto side = [&](int c, int x, int y, int z) {
            return base[((2 * c + (x - x1)) * sc.nz + (z - sc.z0)) * sc.nyp + side_yp(sc, y)];
        };
        if (sc.e_valid) e_at_node_with(side, node[0], node[1], node[2], E);
        else e_at_node(a, f, node[0], node[1], node[2], E);
        b_at_node_with(side, node[0], node[1], node[2], H);
    } else {
        e_at_node(a, f, node[0], node[1], node[2], E);
        b_at_node(a, f, node[0], node[1], node[2], H);
    }
}

// One step of the two-stage EMA of one wavelength (shaders/ntff_accumulate.comp:40-52) on the 24
// phasor values of a point, slot 12 s + 3 q + c.
__device__ __forceinline__ void ntff_update(const float *r, float blend, const float E[3], const float H[3],
                                            float st[24]) {
    const float co[4] = {2.0f * r[0], -2.0f * r[1], 2.0f * r[2], -2.0f * r[3]};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float obs = co[q] * (q < 2 ? E[c] : H[c]);
            float s1 = st[3 * q + c], s2 = st[12 + 3 * q + c];
            s1 += blend * (obs - s1);
            s2 += blend * (s1 - s2);
            st[3 * q + c] = s1;
            st[12 + 3 * q + c] = s2;
        }
}

// The storage format is a template parameter so that the gathers' loads carry no per-load format
// branch and can all be in flight together.
// Workgroups past the far-field points run the port job: one thread per port.
template <int S>
__device__ __forceinline__ bool port_tail(const LatticeArgs &a, const FieldPtrs &f, const NtffBox &box,
                                          const PortJob &pj) {
    const int nb = (box.n_points + 255) / 256;
    if (static_cast<int>(blockIdx.x) < nb) return false;
    const int i = (static_cast<int>(blockIdx.x) - nb) * 256 + static_cast<int>(threadIdx.x);
    if (i < pj.n)
        port_sample(a, f, pj.ports[i], pj.rot, pj.blend, pj.phasors + 8 * i,
                    (pj.rec >= 0 && pj.series) ? pj.series + 2 * (pj.rec * pj.n + i) : nullptr);
    return true;
}

template <int S>
__global__ void __launch_bounds__(256) k_ntff_accumulate(LatticeArgs a, FieldPtrs f, NtffBox box, const float *rot,
                                                         NtffFreqs freqs, float *ph, PortJob pj) {
    f.storage = S;
    if (port_tail<S>(a, f, box, pj)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= box.n_points) return;
    float E[3], H[3];
    ntff_fields(a, f, box, i, E, H);
    const size_t np = static_cast<size_t>(box.n_points);
    for (int fq = 0; fq < freqs.n; ++fq) {
        float *phf = ph + static_cast<size_t>(24 * fq) * np + i;
        float st[24];
#pragma unroll
        for (int j = 0; j < 24; ++j) st[j] = phf[j * np];
        ntff_update(rot + 4 * fq, freqs.blend[fq], E, H, st);
#pragma unroll
        for (int j = 0; j < 24; ++j) phf[j * np] = st[j];
    }
}

// Deferred accumulation over a run of steps: each step only gathers its six field values into a
// sample slot (6 planes of n_points floats); k_ntff_flush then applies the run's updates in step
// order to phasors held in registers. Every phasor goes through the same expressions in the same
// order as with one k_ntff_accumulate per step, so the result is bit-identical, while each phasor
// plane is read and written once per run instead of once per step.
template <int S>
__global__ void __launch_bounds__(256) k_ntff_sample(LatticeArgs a, FieldPtrs f, NtffBox box, float *slot, PortJob pj) {
    f.storage = S;
    if (port_tail<S>(a, f, box, pj)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= box.n_points) return;
    float E[3], H[3];
    ntff_fields(a, f, box, i, E, H);
    const size_t np = static_cast<size_t>(box.n_points);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        __builtin_nontemporal_store(E[c], slot + c * np + i);
        __builtin_nontemporal_store(H[c], slot + (3 + c) * np + i);
    }
}

template <int K>
__global__ void __launch_bounds__(256) k_ntff_flush(NtffBox box, const float *rot, int rot_stride, NtffFreqs freqs,
                                                    const float *samples, float *ph) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= box.n_points) return;
    const size_t np = static_cast<size_t>(box.n_points);
    // the run's samples are read once and serve every wavelength
    float smp[K][6];
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) smp[k][c] = __builtin_nontemporal_load(samples + (6 * k + c) * np + i);
    for (int fq = 0; fq < freqs.n; ++fq) {
        float *phf = ph + static_cast<size_t>(24 * fq) * np + i;
        float st[24];
#pragma unroll
        for (int j = 0; j < 24; ++j) st[j] = phf[j * np];
#pragma unroll
        for (int k = 0; k < K; ++k)
            ntff_update(rot + static_cast<size_t>(k) * rot_stride + 4 * fq, freqs.blend[fq], smp[k], smp[k] + 3, st);
#pragma unroll
        for (int j = 0; j < 24; ++j) phf[j * np] = st[j];
    }
}

// ---------------------------------------------------------------------------------------------
// Flux through axis-aligned rectangles: one wavefront per row of a plane (four rows per
// workgroup), its lanes stepping along the row; each lane sums S_axis = (E x B)_axis at its nodes
// (field_sampling.glsl node values), the wavefront adds its lanes by butterfly shuffles and
// writes one float per row, summed in FP64 on the host (FieldAnalyzer::planeFlux,
// src/FieldAnalyzer.cpp:317-359, defines the quantity). blockIdx.y selects the plane.
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_flux(LatticeArgs a, FieldPtrs f, const FluxPlane *planes, float *partials) {
    const FluxPlane pl = planes[blockIdx.y];
    const int lane = threadIdx.x & 63;
    const int lv = static_cast<int>(blockIdx.x) * 4 + (static_cast<int>(threadIdx.x) >> 6);
    if (lv >= pl.rsize[1]) return;  // whole wavefronts: no barrier follows
    const int vv = pl.rmin[1] + lv;
    float flux = 0.0f;
    for (int lu = lane; lu < pl.rsize[0]; lu += 64) {
        const int uu = pl.rmin[0] + lu;
        int p[3];
        if (pl.axis == 0) { p[0] = pl.index; p[1] = uu; p[2] = vv; }
        else if (pl.axis == 1) { p[0] = uu; p[1] = pl.index; p[2] = vv; }
        else { p[0] = uu; p[1] = vv; p[2] = pl.index; }
        if (p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] < a.nx && p[1] < a.ny && p[2] < a.nz) {
            float e[3], b[3];
            e_at_node(a, f, p[0], p[1], p[2], e);
            b_at_node(a, f, p[0], p[1], p[2], b);
            const int a1 = (pl.axis + 1) % 3, a2 = (pl.axis + 2) % 3;
            flux += e[a1] * b[a2] - e[a2] * b[a1];
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) flux += __shfl_xor(flux, o);
    if (lane == 0) partials[pl.first_group + lv] = flux;
}

// probe_gather.comp (shaders/probe_gather.comp:20-26): 27 E texels then 27 B texels per probe,
// as float4 (x, y, z, alpha) with alpha = cell sigma (E) or eps_r (B).
__global__ void __launch_bounds__(64) k_probe_gather(LatticeArgs a, FieldPtrs f, const int *origins, int n_probes,
                                                     const float *sigma_cell, const float *eps_cell, float *blocks) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n_probes * 54) return;
    const int probe = id / 54, k = id % 54, t = k % 27;
    const int x = min(max(origins[4 * probe + 0] + t % 3, 0), a.nx - 1);
    const int y = min(max(origins[4 * probe + 1] + (t / 3) % 3, 0), a.ny - 1);
    const int z = min(max(origins[4 * probe + 2] + t / 9, 0), a.nz - 1);
    const size_t p = (static_cast<size_t>(z) * a.ny + y) * a.nx + x;
    float *o = blocks + 4 * static_cast<size_t>(id);
    if (k < 27) {
        o[0] = field_load(f, 0, p); o[1] = field_load(f, 1, p); o[2] = field_load(f, 2, p); o[3] = sigma_cell[p];
    } else {
        o[0] = field_load(f, 3, p); o[1] = field_load(f, 4, p); o[2] = field_load(f, 5, p); o[3] = eps_cell[p];
    }
}

// line_accumulate.comp (shaders/line_accumulate.comp:30-48): one thread per quantity sums its
// weighted texel components (clamped indices; 0-2 from E, 3-5 from B), keeps the sum and folds it
// into the two-stage running DFT with the E (kind 0) or B (kind 1) rotation
__global__ void __launch_bounds__(64) k_line_accumulate(LatticeArgs a, FieldPtrs f, const int4 *quantities, int n,
                                                        const int4 *terms, const float *weights, const float *rot,
                                                        float blend, float4 *phasors, float *now) {
    const int q = blockIdx.x * 64 + threadIdx.x;
    if (q >= n) return;
    const int4 info = quantities[q];
    float sum = 0.0f;
    for (int k = info.x; k < info.x + info.y; ++k) {
        const int4 t = terms[k];
        sum += weights[k] * fetch_c(a, f, t.w, t.x, t.y, t.z);
    }
    now[q] = sum;
    const float rc = info.z == 0 ? rot[0] : rot[2], rs = info.z == 0 ? rot[1] : rot[3];
    const float two = 2.0f * sum;
    const float ox = two * rc, oy = two * -rs;
    float4 ph = phasors[q];
    ph.x += blend * (ox - ph.x);
    ph.y += blend * (oy - ph.y);
    ph.z += blend * (ph.x - ph.z);
    ph.w += blend * (ph.y - ph.w);
    phasors[q] = ph;
}

}  // namespace

hipError_t launch_line_accumulate(const LatticeArgs &a, const FieldPtrs &f, const int *quantities, int n,
                                  const int *terms, const float *weights, const float *rot, float blend, float *phasors,
                                  float *now, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    k_line_accumulate<<<(n + 63) / 64, 64, 0, s>>>(a, f, reinterpret_cast<const int4 *>(quantities), n,
                                                   reinterpret_cast<const int4 *>(terms), weights, rot, blend,
                                                   reinterpret_cast<float4 *>(phasors), now);
    return hipGetLastError();
}

hipError_t launch_port_accumulate(const LatticeArgs &a, const FieldPtrs &f, const PortDev *ports, int n_ports,
                                  const float *rot, float blend, float *phasors, float *series, int64_t record_index,
                                  hipStream_t s) {
    if (n_ports <= 0) return hipSuccess;
    const int g = (n_ports + 63) / 64;
    switch (f.storage) {
        case kStorageF32:
            k_port_accumulate<kStorageF32><<<g, 64, 0, s>>>(a, f, ports, n_ports, rot, blend, phasors, series, record_index);
            break;
        case kStorageF16:
            k_port_accumulate<kStorageF16><<<g, 64, 0, s>>>(a, f, ports, n_ports, rot, blend, phasors, series, record_index);
            break;
        case kStorageBF16:
            k_port_accumulate<kStorageBF16><<<g, 64, 0, s>>>(a, f, ports, n_ports, rot, blend, phasors, series,
                                                              record_index);
            break;
        case kStorageF64:
            k_port_accumulate<kStorageF64><<<g, 64, 0, s>>>(a, f, ports, n_ports, rot, blend, phasors, series, record_index);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ntff_accumulate(const LatticeArgs &a, const FieldPtrs &f, const NtffBox &box, const float *rot,
                                  const NtffFreqs &freqs, float *phasors, const PortJob &pj, hipStream_t s) {
    const int g = (box.n_points + 255) / 256 + (pj.n + 255) / 256;
    if (g <= 0) return hipSuccess;
    switch (f.storage) {
        case kStorageF32: k_ntff_accumulate<kStorageF32><<<g, 256, 0, s>>>(a, f, box, rot, freqs, phasors, pj); break;
        case kStorageF16: k_ntff_accumulate<kStorageF16><<<g, 256, 0, s>>>(a, f, box, rot, freqs, phasors, pj); break;
        case kStorageBF16: k_ntff_accumulate<kStorageBF16><<<g, 256, 0, s>>>(a, f, box, rot, freqs, phasors, pj); break;
        case kStorageF64: k_ntff_accumulate<kStorageF64><<<g, 256, 0, s>>>(a, f, box, rot, freqs, phasors, pj); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ntff_sample(const LatticeArgs &a, const FieldPtrs &f, const NtffBox &box, float *slot,
                              const PortJob &pj, hipStream_t s) {
    const int g = (box.n_points + 255) / 256 + (pj.n + 255) / 256;
    if (g <= 0) return hipSuccess;
    switch (f.storage) {
        case kStorageF32: k_ntff_sample<kStorageF32><<<g, 256, 0, s>>>(a, f, box, slot, pj); break;
        case kStorageF16: k_ntff_sample<kStorageF16><<<g, 256, 0, s>>>(a, f, box, slot, pj); break;
        case kStorageBF16: k_ntff_sample<kStorageBF16><<<g, 256, 0, s>>>(a, f, box, slot, pj); break;
        case kStorageF64: k_ntff_sample<kStorageF64><<<g, 256, 0, s>>>(a, f, box, slot, pj); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_ntff_flush(const NtffBox &box, const float *rot, int rot_stride, int steps, const NtffFreqs &freqs,
                             const float *samples, float *phasors, hipStream_t s) {
    if (box.n_points <= 0 || steps <= 0) return hipSuccess;
    static_assert(kNtffRun == 16, "one flush instance per run length");
    const int g = (box.n_points + 255) / 256;
    switch (steps) {
        case 1: k_ntff_flush<1><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 2: k_ntff_flush<2><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 3: k_ntff_flush<3><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 4: k_ntff_flush<4><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 5: k_ntff_flush<5><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 6: k_ntff_flush<6><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 7: k_ntff_flush<7><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 8: k_ntff_flush<8><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 9: k_ntff_flush<9><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 10: k_ntff_flush<10><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 11: k_ntff_flush<11><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 12: k_ntff_flush<12><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 13: k_ntff_flush<13><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 14: k_ntff_flush<14><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 15: k_ntff_flush<15><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        case 16: k_ntff_flush<16><<<g, 256, 0, s>>>(box, rot, rot_stride, freqs, samples, phasors); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_flux(const LatticeArgs &a, const FieldPtrs &f, const FluxPlane *planes, int n_planes,
                       int max_rows, float *partials, hipStream_t s) {
    if (n_planes <= 0 || max_rows <= 0) return hipSuccess;
    dim3 grid((max_rows + 3) / 4, n_planes, 1);
    k_flux<<<grid, 256, 0, s>>>(a, f, planes, partials);
    return hipGetLastError();
}

hipError_t launch_probe_gather(const LatticeArgs &a, const FieldPtrs &f, const int *origins, int n_probes,
                               const float *sigma_cell, const float *eps_cell, float *blocks, hipStream_t s) {
    if (n_probes <= 0) return hipSuccess;
    k_probe_gather<<<(n_probes * 54 + 63) / 64, 64, 0, s>>>(a, f, origins, n_probes, sigma_cell, eps_cell, blocks);
    return hipGetLastError();
}

}  // namespace emfd
