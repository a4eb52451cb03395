// Fused leapfrog in one pass: B^{n+1/2} and E^{n+1} of a whole lattice from one read of E^n and
// B^{n-1/2} and one write of each (ping-pong buffers: src holds the state, dst receives the next).
//
// Work split. A workgroup owns an x segment of 256 cells (one wavefront per row: lane l holds the
// four cells x0 = xs + 4l .. x0 + 3, float4 streams) of NW - 1 consecutive rows y0 .. y0 + NW - 2,
// over one chunk of z planes, which it walks in order (z-march). Wave w holds row y0 - 1 + w: wave
// 0 is the row below the group, whose B^{n+1/2} x, z components the first row's E update needs;
// it computes them and stores nothing. Per plane z:
//   B^{n+1/2}(z) needs E^n(z) and E^n(z+1) of its own row (registers: the march keeps three E
//   planes and two B planes in flight), E^n(y+1) x, z (LDS: published by the wave above one
//   plane earlier; the top wave loads its row above itself) and E^n(x+1) y, z (the next lane,
//   DPP wave shift; lane 63 loads the cell past the segment).
//   E^{n+1}(z) needs B^{n+1/2} of its own row at z and z-1 (registers), of the row below at z
//   (LDS, published before the plane's barrier) and of x-1 y, z (the previous lane; lane 0 of a
//   segment that does not start at x = 0 rebuilds B^{n+1/2}(xs - 1) from its own loads).
// One workgroup barrier per plane: each wave publishes B^{n+1/2}(z) x, z and E^n(z+1) x, z into
// buffers that alternate with the plane's parity. A chunk starts one plane early to rebuild
// B^{n+1/2}(z0 - 1), which it does not store.
//
// Coefficients. Every texel is updated with the coefficients of a vacuum edge: absorber entries
// per depth class staged in LDS per wavefront (as the two-pass kernels stage them), (1, dt)
// outside the absorber; no class ids are read. Texels whose class is not vacuum (materials, feed
// and wire edges) and the E texels next to a thin-wire texel are then recomputed from the same
// src and dst planes by two sparse passes (k_march_fix_b, k_march_fix_e) with faraday() and
// edge_coefficients(), so the result is bit-identical to the two-pass kernels.
#include "port_device.h"
#include "yee_quad.h"

namespace emfd {
namespace {

constexpr int kSeg = 256;  // x cells per wavefront

__device__ __forceinline__ float lane_next(float v) {  // lane i <- lane i + 1 (lane 63 <- 0)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_prev(float v) {  // lane i <- lane i - 1 (lane 0 <- 0)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, true));
}

struct Row4 {
    float v[3][4];
};

// Far-field side columns (SideCols): bit f * 4 + j is set when cell x0 + j of a lane is one of the
// two columns of face f (x = X_f - 1 or X_f). Fixed for the march, so a plane's copy is a test of
// these bits and a few stores on the one or two lanes that hold the columns.
__device__ __forceinline__ uint32_t side_bits4(const SideCols &s, int x0, int nx) {
    uint32_t bits = 0u;
    if (!s.buf) return bits;
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int xo = x0 + j - (s.x[f] - 1);
            if (xo >= 0 && xo <= 1 && x0 + j < nx) bits |= 1u << (f * 4 + j);
        }
    return bits;
}
__device__ __forceinline__ void side_put4(const SideCols &s, uint32_t bits, int comp0, int x0, int y, int z,
                                          const float (&v)[3][4]) {
    const long col = static_cast<long>(s.nz) * s.nyp;
    const long yz = static_cast<long>(z - s.z0) * s.nyp + side_yp(s, y);
#pragma unroll
    for (int f = 0; f < 2; ++f)
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (bits & (1u << (f * 4 + j))) {
                const int xo = x0 + j - (s.x[f] - 1);
                float *d = s.buf + (static_cast<long>(f * 6 + comp0) * 2 + xo) * col + yz;
#pragma unroll
                for (int c = 0; c < 3; ++c) d[c * 2 * col] = v[c][j];
            }
}

// Automatic chunks (kMarchTileZ planes): each wave keeps its row's side column values of the chunk in an LDS record
// [column k = 2 face + (x - X + 1)][component E x y z, B x y z][plane of the chunk], and the
// workgroup stores the records once the chunk is done (side_flush), so that no scattered store
// sits among the march's loads and stores (stores enter the same wait count as the plane loads in
// flight). The column's lane and cell are the same for the whole wave, so a plane's copy is one
// branch on the wave-uniform cell index and six LDS writes on one lane.
template <int J>
__device__ __forceinline__ void side_rec_put(bool on, float *r, const float (&b)[3][4], const float (&e)[3][4]) {
    if (on) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c * kMarchTileZ] = e[c][J];
            r[(3 + c) * kMarchTileZ] = b[c][J];
        }
    }
}
// The workgroup's records as whole 32-byte sectors of the layout (a row group's 7 values and the
// pad slot of one plane, column and component); rows outside the box's range store values the
// gather never reads. (Measured at 768^3 against this: each wave storing its own row's values with
// no barrier, 0.5% slower; a layout keeping a row's chunk contiguous, so that a wave stores 192
// contiguous bytes, 0.5% faster here but 1% slower over the whole step, since the far-field
// gather's lanes read consecutive rows.)
template <int NW>
__device__ __forceinline__ void side_flush(const SideCols &s, const float *recs, int yt, int xs, int xend, int z0,
                                           int z1) {
    const long col = static_cast<long>(s.nz) * s.nyp;
    const int yg = yt - s.gy0;
    // record index sec = (column * 6 + component) * kMarchTileZ + kz; 8 row slots per sector
    for (int k = static_cast<int>(threadIdx.x); k < 2 * 2 * 6 * kMarchTileZ * 8; k += 64 * NW) {
        const int sec = k >> 3, slot = k & 7;
        const int cl = sec / (6 * kMarchTileZ), comp = (sec / kMarchTileZ) % 6, kz = sec % kMarchTileZ;
        const int z = z0 + kz, x = s.x[cl >> 1] - 1 + (cl & 1);
        if (z < z1 && z >= s.z0 && z < s.z0 + s.nz && x >= xs && x < xend) {
            const float v = slot < 7 ? recs[(slot + 1) * (2 * 2 * 6 * kMarchTileZ) + sec] : 0.0f;
            s.buf[(static_cast<long>((cl >> 1) * 6 + comp) * 2 + (cl & 1)) * col + static_cast<long>(z - s.z0) * s.nyp +
                  yg * 8 + slot] = v;
        }
    }
}

// Plane reads go through buffer resources: a byte offset per lane against the plane's 32-bit
// range, where an offset at or past the range (kOob) reads 0 without touching memory. That gives
// the lattice's zero boundaries and idle lanes without a select the compiler could turn into a
// branch (and a wait for every load in flight). Writes use the plane pointer plus the offset.
// Planes are at most 2 GB.
constexpr uint32_t kOob = 0x80000000u;
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc plane_rsrc(const float *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), static_cast<short>(0), static_cast<int>(bytes),
                                             0x00020000);
}
__device__ __forceinline__ void bld4(Rsrc r, uint32_t off, float (&o)[4]) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    o[0] = __uint_as_float(v[0]); o[1] = __uint_as_float(v[1]); o[2] = __uint_as_float(v[2]); o[3] = __uint_as_float(v[3]);
}
__device__ __forceinline__ float bld1(Rsrc r, uint32_t off) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ void gst4(float *p, uint32_t off, const float (&v)[4]) {
    // non-temporal: each value is written once per step
    typedef float f32x4_t __attribute__((ext_vector_type(4)));
    const f32x4_t d = {v[0], v[1], v[2], v[3]};
    __builtin_nontemporal_store(d, reinterpret_cast<f32x4_t *>(reinterpret_cast<char *>(p) + off));
}

// Field storage of the planes: FP32, or binary16 / bfloat16 of value / scale (a power of two per
// component; updates in FP32). ld4 / ld1 return values, st4 stores them, rnd rounds a value
// through the storage (what a store and a load give back), as the 16-bit two-pass kernels do.
template <class T>
__device__ __forceinline__ void ld4(Rsrc r, uint32_t off, float s, float (&o)[4]) {
    if constexpr (sizeof(T) == 4) {
        bld4(r, off, o);
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
        const T *h = reinterpret_cast<const T *>(&v);
#pragma unroll
        for (int j = 0; j < 4; ++j) o[j] = static_cast<float>(h[j]) * s;
    }
}
template <class T>
__device__ __forceinline__ float ld1(Rsrc r, uint32_t off, float s) {
    if constexpr (sizeof(T) == 4) {
        return bld1(r, off);
    } else {
        const auto v = __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
        return static_cast<float>(*reinterpret_cast<const T *>(&v)) * s;
    }
}
template <class T>
__device__ __forceinline__ void st4(float *p, uint32_t off, float inv, const float (&v)[4]) {
    if constexpr (sizeof(T) == 4) {
        gst4(p, off, v);
    } else {
        typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
        u32x2_t raw;
        T *h = reinterpret_cast<T *>(&raw);
#pragma unroll
        for (int j = 0; j < 4; ++j) h[j] = static_cast<T>(v[j] * inv);
        __builtin_nontemporal_store(raw, reinterpret_cast<u32x2_t *>(reinterpret_cast<char *>(p) + off));
    }
}
template <class T>
__device__ __forceinline__ float rnd(float v, float inv, float s) {
    if constexpr (sizeof(T) == 4) return v;
    else return static_cast<float>(static_cast<T>(v * inv)) * s;
}

// LDS layout (floats): xe[2][NW][2][kSeg] | xb[2][NW][2][kSeg] | mat[kMatFloats] |
// side[NW][kSideFloats] | coef[NW][2][K * kCoefStride]
//   xe[p & 1][w]: E^n(row of wave w, plane p) z, x components
//   xb[p & 1][w]: B^{n+1/2}(row of wave w, plane p) x, z components
//   mat: the tile's material classes, entry l (1-based local class) at l * kCoefStride
//   side[w]: far-field side column values of wave w's row over the chunk (automatic chunks)
template <int NW>
__host__ __device__ constexpr size_t march_xch_floats() {
    return static_cast<size_t>(2) * 2 * NW * 2 * kSeg;
}
// one float4 per thread of the 512 copies entries 1 .. 256 (2 per entry)
constexpr int kMatFloats = kMarchCoefStride + 4 * 512;
// [face][column x = X - 1, X][component E x y z, B x y z][plane of the chunk]
constexpr int kSideFloats = 2 * 2 * 6 * kMarchTileZ;
template <int NW>
__host__ __device__ constexpr size_t march_lds_fixed_floats() {
    return march_xch_floats<NW>() + kMatFloats + static_cast<size_t>(NW) * kSideFloats;
}
static_assert(kMarchCoefStride == kCoefStride, "march material entries use the absorber entries' layout");
static_assert(2 * (kMarchMaxLocal + 1) <= 512 + 2, "one float4 copy per thread covers a tile's entries");

// ZC > 0: the z-chunk is ZC planes and the march is unrolled over all ZC + 1 of its planes (the
// leading one included; the first chunk leads with a virtual plane -1 of zero fields, a partial
// last chunk runs on past the lattice without storing), so that the in-flight load slots never
// meet at a loop back-edge, where the compiler would copy them and wait for every load in
// flight. ZC = 0: a loop over any z-chunk.
template <int NW, int ZC, class T>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(16 / NW, 16 / NW))) k_yee_march(LatticeArgs a, FieldPtrs src, FieldPtrs dst, int zchunk,
                                                        int nseg, int nyt, int n_tiles, MarchMat mm) {
    static_assert(ZC == 0 || ZC == kMarchTileZ, "material tables and side records are laid out for the automatic chunks");
    extern __shared__ float lds[];
    const int lane = static_cast<int>(threadIdx.x) & 63;
    const int w = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    // tiles in dispatch order: x segment fastest, then the row group, then the z-chunk, so the
    // workgroups in flight cover one thin z-slab of the whole lattice at a time (handing each XCD
    // a contiguous run of tiles instead measured 4-8% slower at 768^3, and dealing runs of 3 to 48
    // consecutive tiles to one XCD, so that vertically adjacent row groups would share its L2,
    // 1-6% slower)
    const int tile = static_cast<int>(blockIdx.x);
    if (tile >= n_tiles) return;  // uniform over the workgroup, before any barrier
    const int nx = a.nx, ny = a.ny, nz = a.nz, L = a.pml, K = a.pml_k;
    const int seg = tile % nseg;
    const int yt = (tile / nseg) % nyt;
    const int zc = tile / (nseg * nyt);
    const int y = yt * (NW - 1) - 1 + w;
    const int z0 = zc * zchunk;
    const int z1 = min(z0 + zchunk, nz);
    const int zs = ZC > 0 ? z0 - 1 : (z0 > 0 ? z0 - 1 : 0);
    const int xs = seg * kSeg;
    const int x0 = xs + 4 * lane;
    const bool row_ok = y >= 0 && y < ny;
    const bool cell_ok = row_ok && x0 < nx;
    const bool out_row = w > 0 && row_ok;
    const bool top = w == NW - 1;
    const bool up_ok = top && y + 1 < ny;
    const bool left = xs > 0;
    const bool hl_ok = left && lane == 0 && row_ok;               // lane 0 rebuilds B^{n+1/2}(xs - 1)
    const bool hr_ok = xs + kSeg < nx && lane == 63 && row_ok;    // lane 63 reads E^n(xs + 256)
    constexpr uint32_t ES = sizeof(T);  // bytes per stored value
    const uint32_t plane = static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny) * ES;  // bytes per z plane
    const uint32_t yrow = static_cast<uint32_t>(nx) * ES;
    // byte offsets within a plane, kOob for an access outside the lattice
    const uint32_t o_own = cell_ok ? (static_cast<uint32_t>(y) * nx + x0) * ES : kOob;
    const uint32_t o_hl = hl_ok ? (static_cast<uint32_t>(y) * nx + xs - 1) * ES : kOob;
    const uint32_t o_hu = hl_ok && y + 1 < ny ? o_hl + yrow : kOob;
    const uint32_t o_hr = hr_ok ? (static_cast<uint32_t>(y) * nx + xs + kSeg) * ES : kOob;
    const uint32_t o_up = up_ok && x0 < nx ? (static_cast<uint32_t>(y + 1) * nx + x0) * ES : kOob;
    // storage scales of E x y z and B x y z and their inverses (1 for FP32)
    const float sE[3] = {src.scale[0], src.scale[1], src.scale[2]}, sB[3] = {src.scale[3], src.scale[4], src.scale[5]};
    const float iE[3] = {src.inv_scale[0], src.inv_scale[1], src.inv_scale[2]},
                iB[3] = {src.inv_scale[3], src.inv_scale[4], src.inv_scale[5]};
    // The per-lane offsets above stay fixed for the whole march; each load's descriptor covers
    // the one z plane it reads (base + z plane, range one plane), a plane past the chunk's last
    // needed one or past the lattice has range 0 and reads zeros. Every load is issued
    // unconditionally and no address is computed per plane in vector registers, so that the
    // compiler's wait counts stay exact (a skipped load, or a register reused while its load is
    // in flight, makes it wait for every load in flight).
    auto plane_of = [&](const float *p, int z, int zmax, int zmin = 0) {
        const bool ok = z >= zmin && z <= zmax;
        const char *base = reinterpret_cast<const char *>(p) + static_cast<size_t>(ok ? z : 0) * plane;
        return plane_rsrc(reinterpret_cast<const float *>(base), ok ? plane : 0u);
    };

    float *de[3] = {dst.e[0], dst.e[1], dst.e[2]}, *db[3] = {dst.b[0], dst.b[1], dst.b[2]};

    float *xe = lds;
    float *xb = lds + 2 * NW * 2 * kSeg;
    float *mat = lds + march_xch_floats<NW>();
    float *side_l = mat + kMatFloats + w * kSideFloats;
    float *coef = lds + march_lds_fixed_floats<NW>() + static_cast<size_t>(w) * 2 * K * kCoefStride;
    float *coefB = coef, *coefE = coef + K * kCoefStride;
    // The tile's material classes (automatic chunks with tables only): copied into LDS by one
    // unconditional float4 load per thread (a range of 0 reads zeros), issued before the plane
    // loads so that waiting for it waits for nothing else. Per texel local class ids (4 per lane,
    // one byte each) come with the B loads of their plane, from a range of 0 where the tile has
    // no listed class.
    int mat_first = 0, mat_n = 0;
    if (ZC > 0 && mm.tiles) {
        const int2 ti = mm.tiles[tile];
        mat_first = ti.x;
        mat_n = ti.y;
    }
    float mat_v[4];
    {
        const float *cb = mat_n > 0 ? mm.coef + static_cast<size_t>(mat_first) * kMarchCoefStride : src.e[0];
        bld4(plane_rsrc(cb, static_cast<uint32_t>(mat_n) * kMarchCoefStride * 4u), static_cast<uint32_t>(threadIdx.x) * 16u,
             mat_v);
    }
    const uint32_t lid_plane = static_cast<uint32_t>(nx) * static_cast<uint32_t>(ny);
    const uint32_t o_lid = out_row && cell_ok ? static_cast<uint32_t>(y) * nx + x0 : kOob;
    auto lid_of = [&](int z, int zmax) {
        const bool ok = mat_n > 0 && z >= 0 && z <= zmax;
        const char *base = ok ? reinterpret_cast<const char *>(mm.lid) + static_cast<size_t>(z) * lid_plane
                              : reinterpret_cast<const char *>(src.e[0]);
        return plane_rsrc(reinterpret_cast<const float *>(base), ok ? lid_plane : 0u);
    };
    auto xe_at = [&](int p, int wave, int comp) { return xe + ((((p & 1) * NW + wave) * 2 + comp) * kSeg); };
    auto xb_at = [&](int p, int wave, int comp) { return xb + ((((p & 1) * NW + wave) * 2 + comp) * kSeg); };

    // far-field side columns this wave's output row keeps (lanes with column cells, rows in range)
    const uint32_t sbits = ZC == 0 && out_row ? side_bits4(a.side, x0, nx) : 0u;
    const bool side_row = ZC == 0 ? out_row && a.side.buf && y >= a.side.y0 && y < a.side.y0 + a.side.ny &&
                                        __builtin_amdgcn_ballot_w64(sbits != 0u) != 0
                                  : out_row && a.side.buf && y >= a.side.y0 && y < a.side.y0 + a.side.ny;
    const int xend = min(xs + kSeg, nx);
    // automatic chunks: the workgroup stores side columns when one of the columns lies in its
    // segment, its rows meet the box's range and the chunk its z range (uniform)
    bool side_wg = false;
    if (ZC > 0 && a.side.buf) {
        const int ylo = yt * (NW - 1), yhi = ylo + NW - 2;
        bool cols = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int xk = a.side.x[k >> 1] - 1 + (k & 1);
            cols = cols || (xk >= xs && xk < xend);
        }
        side_wg = cols && yhi >= a.side.y0 && ylo < a.side.y0 + a.side.ny && z0 < a.side.z0 + a.side.nz &&
                  z1 > a.side.z0;
    }
    // depth class per cell of the lane (0 outside the lattice: any finite entry will do there)
    int cx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) cx[j] = x0 + j < nx ? pml_class(x0 + j, nx, L) : 0;
    const int cxl = left ? pml_class(xs - 1, nx, L) : 0;

    // E^n of the row at plane z, and lane 0's E^n x, y, z at xs - 1
    const int ze_max = min(z1, nz - 1), zb_max = z1 - 1;  // last E and B^n planes the chunk reads
    // Values only B^{n+1/2} z depends on are not read on the leading plane (its B z is never
    // used: no E is updated there), nor is E^n z on the plane past the chunk (only B x, y of the
    // last plane read that plane): those loads get a range of 0 and move no bytes.
    // Wave 0 (the row below the group) only feeds B^{n+1/2} x, z of its row on the chunk's own
    // planes to the first row's E update: it reads no B y, nothing on the leading plane, and of
    // the plane past the chunk only E y (B x of the last plane).
    const bool halo = w == 0;  // wave-uniform
    const int zlo = halo ? z0 : 0, zno = -2;  // zno: a last plane below every plane, range 0
    auto load_e = [&](int z, Row4 &r, float (&hl)[3]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const Rsrc q = plane_of(src.e[c], z, c == 2 || (c == 0 && halo) ? zb_max : ze_max, zlo);
            ld4<T>(q, o_own, sE[c], r.v[c]);
            hl[c] = ld1<T>(q, o_hl, sE[c]);
        }
    };
    struct BSet {
        Row4 b;
        float hb[2];     // lane 0: B^n y, z at xs - 1
        float hu;        // lane 0: E^n x at (xs - 1, y + 1)
        float hr[2];     // lane 63: E^n y, z at xs + 256
        float eu[2][4];  // top wave: E^n z, x of row y + 1
        uint32_t lid;    // local material classes of the lane's four cells (byte j: cell j)
    };
    auto load_b = [&](int z, BSet &s) {
        const Rsrc b0 = plane_of(src.b[0], z, zb_max, zlo), b1 = plane_of(src.b[1], z, halo ? zno : zb_max),
                   b2 = plane_of(src.b[2], z, zb_max, z0);
        // e0 (lane 0's E x at y + 1, the top wave's E x of row y + 1) and e1 (lane 63's E y past
        // the segment) feed B z only; wave 0 reads e1 alone (lane 0's values and the top wave's
        // feed only the E update, which wave 0 does not make)
        const Rsrc e0 = plane_of(src.e[0], z, halo ? zno : zb_max, z0), e1 = plane_of(src.e[1], z, zb_max, z0),
                   e2 = plane_of(src.e[2], z, halo ? zno : zb_max);
        ld4<T>(b0, o_own, sB[0], s.b.v[0]);
        ld4<T>(b1, o_own, sB[1], s.b.v[1]);
        ld4<T>(b2, o_own, sB[2], s.b.v[2]);
        s.hb[0] = ld1<T>(b1, o_hl, sB[1]);
        s.hb[1] = ld1<T>(b2, o_hl, sB[2]);
        s.hu = ld1<T>(e0, o_hu, sE[0]);
        s.hr[0] = ld1<T>(e1, o_hr, sE[1]);
        s.hr[1] = ld1<T>(e2, o_hr, sE[2]);
        ld4<T>(e2, o_up, sE[2], s.eu[0]);  // kOob but for the top wave
        ld4<T>(e0, o_up, sE[0], s.eu[1]);
        s.lid = __builtin_amdgcn_raw_buffer_load_b32(lid_of(z, zb_max), o_lid, 0, 0);
    };
    // absorber coefficients of this wave's row at plane z: entry cx for (cx, class(y), class(z));
    // the wave issues the table loads, waits, and stores them (one entry per lane, more when the
    // absorber is thicker than 31 cells)
    auto stage = [&](int z) {
        const int yy = row_ok ? y : 0;
        // planes outside the lattice (the virtual plane -1, the planes a partial last chunk runs
        // on past nz - 1) store nothing: their entries come from the nearest lattice plane, since
        // their own depth class would index past the absorber table
        const int zz = min(max(z, 0), nz - 1);
        const CoefLoad cb = coef_load<1>(a, lane, yy, zz, kPmlBa);
        const CoefLoad ce = coef_load<1>(a, lane, yy, zz, kPmlEa);
        if (lane < K) {
            coef_store(a, coefB, lane, cb);
            coef_store(a, coefE, lane, ce);
        }
        for (int i = lane + 64; i < K; i += 64) {
            coef_store(a, coefB, i, coef_load<1>(a, i, yy, zz, kPmlBa));
            coef_store(a, coefE, i, coef_load<1>(a, i, yy, zz, kPmlEa));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // rings: E^n planes z, z + 1 (ec, en) and two in flight (P, Q: z + 2, z + 3); B^n plane z (bc)
    // and two in flight (Pb, Qb: z + 1, z + 2). The loop is unrolled by two so that the in-flight
    // slots trade names instead of being copied: a slot is read (copied into en / bc) one plane
    // after its loads were issued, never in the iteration that issued them.
    Row4 ec, en, eP, eQ;
    float hlc[3], hln[3], hlP[3], hlQ[3];
    BSet bc, bP, bQ;
    float bm[2][4];  // B^{n+1/2}(z - 1) x, y of the row
#pragma unroll
    for (int j = 0; j < 4; ++j) bm[0][j] = bm[1][j] = 0.0f;

    load_e(zs, ec, hlc);
    load_e(zs + 1, en, hln);
    load_e(zs + 2, eP, hlP);
    load_b(zs, bc);
    load_b(zs + 1, bP);
    *reinterpret_cast<float4 *>(mat + kMarchCoefStride + 4 * threadIdx.x) = make_float4(mat_v[0], mat_v[1], mat_v[2], mat_v[3]);
    // publish E^n(zs) z, x of the row for the wave below
    *reinterpret_cast<float4 *>(xe_at(zs, w, 0) + 4 * lane) = make_float4(ec.v[2][0], ec.v[2][1], ec.v[2][2], ec.v[2][3]);
    *reinterpret_cast<float4 *>(xe_at(zs, w, 1) + 4 * lane) = make_float4(ec.v[0][0], ec.v[0][1], ec.v[0][2], ec.v[0][3]);
    __syncthreads();

    int cz_prev = -1;
    auto plane_step = [&](int z, Row4 &P, Row4 &Q, float (&hP)[3], float (&hQ)[3], BSet &Pb, BSet &Qb) {
        const int czz = pml_class(min(max(z, 0), nz - 1), nz, L);
        if (czz != cz_prev) {  // uniform over the workgroup
            stage(z);
            cz_prev = czz;
        }
        // prefetch E^n(z + 3) and B^n(z + 2): two planes ahead of their use
        load_e(z + 3, Q, hQ);
        load_b(z + 2, Qb);

        // ---- B^{n+1/2}(z) ----
        // E^n(y + 1, z) z, x: the wave above's row from LDS, the top wave's own loads (a select,
        // not a branch: the loaded registers stay in use every plane, so they are never reused
        // while their loads are in flight)
        float eyu[4], exu[4];
        {
            const int wu = top ? w : w + 1;
            const float4 u0 = *reinterpret_cast<const float4 *>(xe_at(z, wu, 0) + 4 * lane);
            const float4 u1 = *reinterpret_cast<const float4 *>(xe_at(z, wu, 1) + 4 * lane);
            const float l0[4] = {u0.x, u0.y, u0.z, u0.w}, l1[4] = {u1.x, u1.y, u1.z, u1.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                eyu[j] = top ? bc.eu[0][j] : l0[j];
                exu[j] = top ? bc.eu[1][j] : l1[j];
            }
        }
        // E^n(x + 1) y, z for the lane's last cell: the next lane's first, or past the segment
        float eyx = lane_next(ec.v[1][0]), ezx = lane_next(ec.v[2][0]);
        if (lane == 63) { eyx = bc.hr[0]; ezx = bc.hr[1]; }
        float bn[3][4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float e0x = ec.v[0][j], e0y = ec.v[1][j], e0z = ec.v[2][j];
            const float exn_y = j < 3 ? ec.v[1][j + 1] : eyx, exn_z = j < 3 ? ec.v[2][j + 1] : ezx;
            // shaders/update_B.comp:41-47 term order
            const float t0x = eyu[j] - e0z, t0y = en.v[0][j] - e0x, t0z = exn_y - e0y;
            const float t1x = en.v[1][j] - e0y, t1y = exn_z - e0z, t1z = exu[j] - e0x;
            const float cxv = t0x - t1x, cyv = t0y - t1y, czv = t0z - t1z;
            float ba[3], bb[3];
            read_coef(coefB + cx[j] * kCoefStride, ba, bb);
            bn[0][j] = ba[0] * bc.b.v[0][j] - bb[0] * cxv;
            bn[1][j] = ba[1] * bc.b.v[1][j] - bb[1] * cyv;
            bn[2][j] = ba[2] * bc.b.v[2][j] - bb[2] * czv;
        }
        if (!row_ok || z < 0) {
            // a row outside the lattice (the group below row 0) or the virtual plane -1: B there is
            // the boundary's 0, which the first row's or plane's E update reads
            // (shaders/update_E.comp:27-29)
#pragma unroll
            for (int j = 0; j < 4; ++j) bn[0][j] = bn[1][j] = bn[2][j] = 0.0f;
        }
        // the E update reads B^{n+1/2} as stored (16-bit storage: rounded)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) bn[c][j] = rnd<T>(bn[c][j], iB[c], sB[c]);
        // lane 0 of a segment past x = 0: B^{n+1/2}(xs - 1) y, z from its own loads
        float hby = 0.0f, hbz = 0.0f;
        if (left && lane == 0) {
            float ba[3], bb[3];
            read_coef(coefB + cxl * kCoefStride, ba, bb);
            const float t0y = hln[0] - hlc[0], t1y = ec.v[2][0] - hlc[2];
            const float t0z = ec.v[1][0] - hlc[1], t1z = bc.hu - hlc[0];
            hby = rnd<T>(ba[1] * bc.hb[0] - bb[1] * (t0y - t1y), iB[1], sB[1]);
            hbz = rnd<T>(ba[2] * bc.hb[1] - bb[2] * (t0z - t1z), iB[2], sB[2]);
        }
        // publish B^{n+1/2}(z) x, z and E^n(z + 1) z, x
        *reinterpret_cast<float4 *>(xb_at(z, w, 0) + 4 * lane) = make_float4(bn[0][0], bn[0][1], bn[0][2], bn[0][3]);
        *reinterpret_cast<float4 *>(xb_at(z, w, 1) + 4 * lane) = make_float4(bn[2][0], bn[2][1], bn[2][2], bn[2][3]);
        *reinterpret_cast<float4 *>(xe_at(z + 1, w, 0) + 4 * lane) =
            make_float4(en.v[2][0], en.v[2][1], en.v[2][2], en.v[2][3]);
        *reinterpret_cast<float4 *>(xe_at(z + 1, w, 1) + 4 * lane) =
            make_float4(en.v[0][0], en.v[0][1], en.v[0][2], en.v[0][3]);
        // B^{n+1/2}(x - 1) y, z for the lane's first cell (all lanes take part in the shift)
        float byx = lane_prev(bn[1][3]), bzx = lane_prev(bn[2][3]);
        if (lane == 0) { byx = hby; bzx = hbz; }
        __syncthreads();

        // ---- E^{n+1}(z) ----
        if (out_row && z >= z0 && z < z1) {
            const float4 d0 = *reinterpret_cast<const float4 *>(xb_at(z, w - 1, 0) + 4 * lane);
            const float4 d1 = *reinterpret_cast<const float4 *>(xb_at(z, w - 1, 1) + 4 * lane);
            const float bxd[4] = {d0.x, d0.y, d0.z, d0.w}, bzd[4] = {d1.x, d1.y, d1.z, d1.w};
            float e[3][4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float bxm_y = j > 0 ? bn[1][j - 1] : byx, bxm_z = j > 0 ? bn[2][j - 1] : bzx;
                // shaders/update_E.comp:35-40 curl order
                const float curl0 = (bn[2][j] - bzd[j]) - (bn[1][j] - bm[1][j]);
                const float curl1 = (bn[0][j] - bm[0][j]) - (bn[2][j] - bxm_z);
                const float curl2 = (bn[1][j] - bxm_y) - (bn[0][j] - bxd[j]);
                float ca[3], cb[3];
                // a listed material class, else the vacuum edge's (absorber) coefficients
                const uint32_t l = (bc.lid >> (8 * j)) & 0xffu;
                read_coef(l != 0u ? mat + l * kMarchCoefStride : coefE + cx[j] * kCoefStride, ca, cb);
                e[0][j] = rnd<T>(ca[0] * ec.v[0][j] + cb[0] * curl0, iE[0], sE[0]);
                e[1][j] = rnd<T>(ca[1] * ec.v[1][j] + cb[1] * curl1, iE[1], sE[1]);
                e[2][j] = rnd<T>(ca[2] * ec.v[2][j] + cb[2] * curl2, iE[2], sE[2]);
            }
            const uint32_t o = o_own + static_cast<uint32_t>(z) * plane;
            if (cell_ok) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    st4<T>(db[c], o, iB[c], bn[c]);
                    st4<T>(de[c], o, iE[c], e[c]);
                }
            }
            if constexpr (ZC > 0) {
                // into the wave's LDS record; stored once the chunk is done
                if (side_row) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        const int d = a.side.x[k >> 1] - 1 + (k & 1) - xs;  // uniform
                        if (d >= 0 && xs + d < xend) {
                            const bool on = lane == (d >> 2);
                            float *r = side_l + k * 6 * kMarchTileZ + (z - z0);
                            switch (d & 3) {
                                case 0: side_rec_put<0>(on, r, bn, e); break;
                                case 1: side_rec_put<1>(on, r, bn, e); break;
                                case 2: side_rec_put<2>(on, r, bn, e); break;
                                default: side_rec_put<3>(on, r, bn, e); break;
                            }
                        }
                    }
                }
            } else if (side_row && z >= a.side.z0 && z < a.side.z0 + a.side.nz && sbits != 0u) {
                side_put4(a.side, sbits, 3, x0, y, z, bn);
                side_put4(a.side, sbits, 0, x0, y, z, e);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) { bm[0][j] = bn[0][j]; bm[1][j] = bn[1][j]; }
        ec = en;
        en = P;
#pragma unroll
        for (int c = 0; c < 3; ++c) { hlc[c] = hln[c]; hln[c] = hP[c]; }
        bc = Pb;
    };
    if constexpr (ZC > 0) {
#pragma unroll
        for (int k = 0; k <= ZC; k += 2) {
            plane_step(zs + k, eP, eQ, hlP, hlQ, bP, bQ);
            if (k + 1 <= ZC) plane_step(zs + k + 1, eQ, eP, hlQ, hlP, bQ, bP);
        }
        if (side_wg) {
            __syncthreads();
            side_flush<NW>(a.side, mat + kMatFloats, yt, xs, xend, z0, z1);
        }
    } else {
        for (int z = zs; z < z1; z += 2) {
            plane_step(z, eP, eQ, hlP, hlQ, bP, bQ);
            if (z + 1 < z1) plane_step(z + 1, eQ, eP, hlQ, hlP, bQ, bP);
        }
    }
}

// List item of this thread. Workgroups are dealt to the 8 XCDs in turn; each XCD is given one
// contiguous eighth of the (sorted) list instead, so that a texel's -y and -z neighbours, read a
// row or a plane of the list earlier, were read through the same XCD's L2.
__device__ __forceinline__ int fix_item(int n) {
    const int blocks = static_cast<int>(gridDim.x), per = blocks / 8;  // the grid is a multiple of 8
    const int b = static_cast<int>(blockIdx.x);
    return ((b % 8) * per + b / 8) * 256 + static_cast<int>(threadIdx.x);
}

// B^{n+1/2} of the listed texels (thin-wire classes), from src as update_B.comp computes it
// (faraday(): wire factors and absorber from the class and absorber tables).
__global__ void __launch_bounds__(256) k_march_fix_b(LatticeArgs a, FieldPtrs src, FieldPtrs dst,
                                                     const uint32_t *list, int n) {
    const int i = fix_item(n);
    if (i >= n) return;
    const uint32_t q = list[i], row = q / static_cast<uint32_t>(a.nx);  // at most 2^29 texels
    const size_t p = q;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int x = static_cast<int>(q - row * static_cast<uint32_t>(nx)), y = static_cast<int>(row % static_cast<uint32_t>(ny)),
              z = static_cast<int>(row / static_cast<uint32_t>(ny));
    const size_t sy = nx, sz = static_cast<size_t>(nx) * ny;
    float e0[3], exn[3], eyn[3], ezn[3], b[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e0[c] = field_load(src, c, p);
        exn[c] = x + 1 < nx ? field_load(src, c, p + 1) : 0.0f;
        eyn[c] = y + 1 < ny ? field_load(src, c, p + sy) : 0.0f;
        ezn[c] = z + 1 < nz ? field_load(src, c, p + sz) : 0.0f;
        b[c] = field_load(src, 3 + c, p);
    }
    faraday(a, x, y, z, e0, exn, eyn, ezn, load_id(a, p), b);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        field_store(dst, 3 + c, p, b[c]);
        b[c] = storage_value(dst, 3 + c, b[c]);
    }
    if (a.side.buf) {
        for (int face = 0; face < 2; ++face) {
            const long k = side_index(a.side, face, 3, x, y, z);
            if (k < 0) continue;
            const long plane = 2L * a.side.nz * a.side.nyp;
            for (int c = 0; c < 3; ++c) a.side.buf[k + c * plane] = b[c];
        }
    }
}

// E^{n+1} of the listed texels (classes other than vacuum, and the E texels whose curl reads a
// thin-wire texel's B), from src E and dst B as update_E.comp computes it (edge_coefficients()).
__global__ void __launch_bounds__(256) k_march_fix_e(LatticeArgs a, FieldPtrs src, FieldPtrs dst,
                                                     const uint32_t *list, int n) {
    const int i = fix_item(n);
    if (i >= n) return;
    const uint32_t q = list[i], row = q / static_cast<uint32_t>(a.nx);
    const size_t p = q;
    const int nx = a.nx, ny = a.ny;
    const int x = static_cast<int>(q - row * static_cast<uint32_t>(nx)), y = static_cast<int>(row % static_cast<uint32_t>(ny)),
              z = static_cast<int>(row / static_cast<uint32_t>(ny));
    const size_t sy = nx, sz = static_cast<size_t>(nx) * ny;
    float b0[3], bX[3], bY[3], bZ[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        b0[c] = field_load(dst, 3 + c, p);
        bX[c] = x > 0 ? field_load(dst, 3 + c, p - 1) : 0.0f;
        bY[c] = y > 0 ? field_load(dst, 3 + c, p - sy) : 0.0f;
        bZ[c] = z > 0 ? field_load(dst, 3 + c, p - sz) : 0.0f;
    }
    const float curl[3] = {(b0[2] - bY[2]) - (b0[1] - bZ[1]), (b0[0] - bZ[0]) - (b0[2] - bX[2]),
                           (b0[1] - bX[1]) - (b0[0] - bY[0])};
    float ca[3], cb[3];
    edge_coefficients<true>(a, x, y, z, load_id(a, p), ca, cb);
    float e[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        e[c] = ca[c] * field_load(src, c, p) + cb[c] * curl[c];
        field_store(dst, c, p, e[c]);
        e[c] = storage_value(dst, c, e[c]);  // the far-field columns keep what the plane holds
    }
    if (a.side.buf) {
        for (int face = 0; face < 2; ++face) {
            const long k = side_index(a.side, face, 0, x, y, z);
            if (k < 0) continue;
            const long plane = 2L * a.side.nz * a.side.nyp;
            for (int c = 0; c < 3; ++c) a.side.buf[k + c * plane] = e[c];
        }
    }
}

template <int NW, class T>
hipError_t launch_nw(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int zchunk, const MarchMat &mm,
                     hipStream_t s) {
    const int nseg = (a.nx + kSeg - 1) / kSeg;
    const int nyt = (a.ny + (NW - 2)) / (NW - 1);
    if (zchunk <= 0) {
        // thin z-chunks: the workgroups in flight then cover a thin slab, which streams best
        // although each chunk re-reads a leading plane (768^3 vacuum, unrolled marches: chunks of
        // 4 planes 3.89 ms per step, 6 planes 3.79, 8 planes 3.82, 12 planes 3.86, 16 planes 3.91;
        // with a loop, 48 planes 4.00, 96 planes 4.21; with the material tables and side records,
        // 6 planes measured 3.82 against 3.84 ms in vacuum but 3.97 against 3.94 ms in the 768^3
        // precision scene, whose per-chunk table loads and side stores then come more often)
        zchunk = kMarchTileZ;
    }
    if (zchunk > a.nz) zchunk = a.nz;
    const int nzc = (a.nz + zchunk - 1) / zchunk;
    const int n_tiles = nseg * nyt * nzc;
    const size_t lds = (march_lds_fixed_floats<NW>() + static_cast<size_t>(NW) * 2 * a.pml_k * kCoefStride) * sizeof(float);
    auto allow_lds = [&](const void *kernel, size_t &raised) -> hipError_t {
        // dynamic LDS past 64 KB must be allowed per kernel
        if (lds <= 65536 || lds <= raised) return hipSuccess;
        const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 static_cast<int>(lds));
        if (e == hipSuccess) raised = lds;
        return e;
    };
    if (zchunk == kMarchTileZ) {
        static size_t raised = 0;
        if (hipError_t e = allow_lds(reinterpret_cast<const void *>(&k_yee_march<NW, kMarchTileZ, T>), raised); e != hipSuccess)
            return e;
        k_yee_march<NW, kMarchTileZ, T><<<n_tiles, 64 * NW, lds, s>>>(a, src, dst, zchunk, nseg, nyt, n_tiles, mm);
    } else {
        static size_t raised = 0;
        if (hipError_t e = allow_lds(reinterpret_cast<const void *>(&k_yee_march<NW, 0, T>), raised); e != hipSuccess)
            return e;
        k_yee_march<NW, 0, T><<<n_tiles, 64 * NW, lds, s>>>(a, src, dst, zchunk, nseg, nyt, n_tiles, MarchMat{});
    }
    return hipGetLastError();
}

}  // namespace

int yee_march_supported(const LatticeArgs &a) {
    const size_t bytes = static_cast<size_t>(a.nx) * a.ny * a.nz * 4;
    return a.nx % 4 == 0 && bytes <= kOob ? 1 : 0;
}

hipError_t launch_yee_march(const LatticeArgs &a, const FieldPtrs &src, const FieldPtrs &dst, int cfg,
                            const uint32_t *fix_b, int n_fix_b, const uint32_t *fix_e, int n_fix_e,
                            const MarchMat &mm, const uint32_t *fix_e_rest, int n_fix_e_rest,
                            const MarchCpml *cpml, hipStream_t s) {
    if (!yee_march_supported(a)) return hipErrorInvalidValue;
    const int zchunk = cfg;  // 0: automatic
    // the material tables are laid out for the automatic chunks
    const bool tables = mm.lid && mm.tiles && mm.coef && (zchunk == 0 || zchunk == kMarchTileZ) && a.nz >= kMarchTileZ;
    const MarchMat use = tables ? mm : MarchMat{};
    if (tables) {
        fix_e = fix_e_rest;
        n_fix_e = n_fix_e_rest;
    }
    hipError_t e = src.storage == kStorageF16    ? launch_nw<8, _Float16>(a, src, dst, zchunk, use, s)
                   : src.storage == kStorageBF16 ? launch_nw<8, __bf16>(a, src, dst, zchunk, use, s)
                   : src.storage == kStorageF32  ? launch_nw<8, float>(a, src, dst, zchunk, use, s)
                                                 : hipErrorInvalidValue;
    if (e != hipSuccess) return e;
    // (one launch for both lists, the E items recomputing the thin-wire B they read, measured no
    // faster in the 768^3 precision scene: the B pass is a few microseconds)
    auto grid8 = [](int n) { return ((n + 255) / 256 + 7) / 8 * 8; };
    if (n_fix_b > 0) {
        k_march_fix_b<<<grid8(n_fix_b), 256, 0, s>>>(a, src, dst, fix_b, n_fix_b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (cpml && (e = launch_cpml_march(a, src, dst, cpml->coef, cpml->psi, false, s)) != hipSuccess) return e;
    if (n_fix_e > 0) {
        k_march_fix_e<<<grid8(n_fix_e), 256, 0, s>>>(a, src, dst, fix_e, n_fix_e);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (cpml && (e = launch_cpml_march(a, src, dst, cpml->coef, cpml->psi, true, s)) != hipSuccess) return e;
    return hipSuccess;
}

}  // namespace emfd
