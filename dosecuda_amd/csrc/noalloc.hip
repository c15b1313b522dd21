// noalloc.hip -- proton WET (ray trace + lateral smoothing) with no device
// allocation on any call: the quad-packed padded volume and the unsmoothed
// WET live in one fixed file-scope device array, so callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc) can still compute a field's WET.  Same kernels and result
// as dh_proton_raytrace_dev; only where the scratch comes from differs.
//
// Each GPU holds its own copy of the array from the library's first use on
// that device until the process ends.  Every call on one GPU shares it, so
// calls that overlap on one GPU (different streams, or threads) race; calls
// ordered on one stream do not.  The array is fully rewritten by each call
// (padding kernel, then the ray trace writes every voxel), so nothing needs
// clearing between calls.
#include "dh_runtime.h"

#include <mutex>

namespace {
constexpr size_t kScratchBytes = (size_t)64 << 20;     // 64 MiB per device
constexpr int kMaxDev = 64;
std::mutex g_mu;
void *g_scratch[kMaxDev];                              // device address per GPU
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Resolves the array's device address on `dev` (which must be current) once.
void dh_noalloc_setup(int dev)
{
    DH_REQUIRE(dev >= 0 && dev < kMaxDev, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[dev]) return;
    void *p = nullptr;
    DH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DH_REQUIRE(p != nullptr, "no device address for the fixed scratch array");
    g_scratch[dev] = p;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

extern "C" int dh_proton_raytrace_dev_noalloc(int gpu_id, const dh_grid *grid,
                                              const dh_proton_beam *beam, const float *d_rlsp,
                                              float *d_wet_out, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(grid && beam && d_rlsp && d_wet_out, "null argument");
        int n_dev = 0;
        DH_HIP(hipGetDeviceCount(&n_dev));
        if (gpu_id < 0 || gpu_id >= n_dev || gpu_id >= kMaxDev)
            throw dh_error(DH_E_RUNTIME, hipGetErrorString(hipErrorInvalidDevice));
        const dh_geom g = dh_proton_geom(grid, beam);   // validates the grid (host only)
        const dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
        const size_t pad_bytes = align256(sizeof(float4) * (size_t)p.pxy * p.pz);
        const size_t raw_bytes = align256(sizeof(float) * (size_t)g.nx * g.ny * g.nz);
        if (pad_bytes + raw_bytes > kScratchBytes) {
            throw dh_error(DH_E_INVALID,
                           "volume too large for the no-allocation WET: needs " +
                               std::to_string(pad_bytes + raw_bytes) +
                               " bytes of scratch, the fixed limit is 67108864 (64 MiB; "
                               "(nx+3)(ny+3)(nz+3)*16 + nx*ny*nz*4 bytes, each rounded up to 256)");
        }
        DH_HIP(hipSetDevice(gpu_id));
        dh_noalloc_setup(gpu_id);
        unsigned char *base = (unsigned char *)g_scratch[gpu_id];
        float4 *vpad = (float4 *)base;
        float *raw = (float *)(base + pad_bytes);
        hipStream_t s = (hipStream_t)stream;
        dh_launch_pad_quad(d_rlsp, vpad, g.nx, g.ny, g.nz, 0.0f, s);   // IMPTClasses.cu:292, border 0
        dh_launch_raytrace(g, vpad, raw, s);
        dh_launch_smooth(g, raw, d_wet_out, s);
        DH_HIP(hipGetLastError());
    });
}
