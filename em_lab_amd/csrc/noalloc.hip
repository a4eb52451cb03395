// Device memory of the no-allocation flux measurement (emfd_flux_measure_noalloc): one file-scope
// array per GPU, holding the plane list, one partial per plane row and the per-surface sums, so
// that a measurement needs no device allocation. Each GPU's copy lives until the process ends and
// keeps the last call's contents; calls that overlap on one GPU race on it.
#include "context.h"

#include <mutex>

namespace emfd {
namespace {

constexpr size_t kPlanesOff = 0;
constexpr size_t kRefsOff = 16384;
constexpr size_t kSumsOff = 20480;
constexpr size_t kPartialsOff = 24576;
constexpr size_t kScratchBytes = kPartialsOff + sizeof(float) * static_cast<size_t>(kFluxScratchRows);
static_assert(sizeof(FluxPlane) * 6 * kFluxScratchSurfaces <= kRefsOff - kPlanesOff, "plane region");
static_assert(sizeof(FluxRef) * 6 * kFluxScratchSurfaces <= kSumsOff - kRefsOff, "ref region");
static_assert(sizeof(double) * kFluxScratchSurfaces <= kPartialsOff - kSumsOff, "sum region");

// 1 MiB + 24 KiB per GPU. On the host this name is only a handle for hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
unsigned char *g_ptr[kMaxDevices] = {};

// One thread per surface: its planes in list order, each an FP64 sum over its rows, signed into
// the surface's sum (the same operations, in the same order, as the host reduction of
// emfd_flux_measure).
__global__ void __launch_bounds__(64) k_flux_surfaces(const FluxPlane *planes, const FluxRef *refs, int n_planes,
                                                      const float *partials, int n_surfaces, double *sums) {
    const int s = static_cast<int>(blockIdx.x) * 64 + static_cast<int>(threadIdx.x);
    if (s >= n_surfaces) return;
    double acc = 0.0;
    for (int i = 0; i < n_planes; ++i) {
        if (refs[i].surface != s) continue;
        double sum = 0.0;
        const float *p = partials + planes[i].first_group;
        for (int k = 0; k < planes[i].rsize[1]; ++k) sum += static_cast<double>(p[k]);
        acc += static_cast<double>(refs[i].sign) * sum;
    }
    sums[s] = acc;
}

}  // namespace

int flux_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(EMFD_ERR_INVALID, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_ptr[device]) return EMFD_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress");
    g_ptr[device] = static_cast<unsigned char *>(p);
    return EMFD_OK;
}

int flux_scratch(int device, FluxScratch *out) {
    if (device < 0 || device >= kMaxDevices) return fail(EMFD_ERR_INVALID, "device index beyond 64");
    unsigned char *base;
    {
        std::lock_guard<std::mutex> lock(g_mutex);
        base = g_ptr[device];
    }
    if (!base) return fail(EMFD_ERR_STATE, "no-allocation scratch not set up on this device");
    out->planes = reinterpret_cast<FluxPlane *>(base + kPlanesOff);
    out->refs = reinterpret_cast<FluxRef *>(base + kRefsOff);
    out->sums = reinterpret_cast<double *>(base + kSumsOff);
    out->partials = reinterpret_cast<float *>(base + kPartialsOff);
    return EMFD_OK;
}

hipError_t launch_flux_surfaces(const FluxPlane *planes, const FluxRef *refs, int n_planes, const float *partials,
                                int n_surfaces, double *sums, hipStream_t s) {
    if (n_surfaces <= 0) return hipSuccess;
    k_flux_surfaces<<<(n_surfaces + 63) / 64, 64, 0, s>>>(planes, refs, n_planes, partials, n_surfaces, sums);
    return hipGetLastError();
}

}  // namespace emfd
