// tuner.hip -- autotuner helpers and the ray-trace / smoothing tuner.
// A tuner runs every candidate variant of a kernel on the caller's data,
// times it with HIP events (median of reps launches after one warm-up),
// requires its output to equal candidate 0's bit for bit (the variants
// only remap threads to voxels or move a table between LDS and global
// memory, so any difference is a defect), and stores the fastest choice in
// the tuning table under the shape's key, where the launchers find it.
//
// No reference counterpart: DoseCUDA's launch shapes are compile-time
// constants (CudaClasses.cuh:14).
#include "dh_runtime.h"

#include <algorithm>
#include <vector>

namespace {

__global__ void k_count_diff(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, int64_t n,
                             unsigned long long *__restrict__ out)
{
    unsigned long long c = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        c += a[i] != b[i];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (__lane_id() == 0 && c) atomicAdd(out, c);
}

}  // namespace

int64_t dh_count_diff(const void *a, const void *b, int64_t n_words, hipStream_t s)
{
    unsigned long long *d = (unsigned long long *)dh_scratch(dh_current_device(), DH_SLOT_COUNTER, 256);
    unsigned long long h = 0;
    DH_HIP(hipMemsetAsync(d, 0, sizeof(h), s));     // tuners never run under stream capture
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n_words + 255) / 256, 2048));
    hipLaunchKernelGGL(k_count_diff, dim3(grid), dim3(256), 0, s, (const uint32_t *)a, (const uint32_t *)b, n_words,
                       d);
    DH_HIP(hipGetLastError());
    DH_HIP(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    DH_HIP(hipStreamSynchronize(s));
    return (int64_t)h;
}

float dh_time_median(const std::function<void()> &f, int reps, hipStream_t s)
{
    reps = std::max(1, reps);
    f();
    std::vector<hipEvent_t> ev(2 * reps);
    for (auto &e : ev) DH_HIP(hipEventCreate(&e));
    for (int r = 0; r < reps; r++) {
        DH_HIP(hipEventRecord(ev[2 * r], s));
        f();
        DH_HIP(hipEventRecord(ev[2 * r + 1], s));
    }
    DH_HIP(hipStreamSynchronize(s));
    std::vector<float> ms(reps);
    for (int r = 0; r < reps; r++) DH_HIP(hipEventElapsedTime(&ms[r], ev[2 * r], ev[2 * r + 1]));
    for (auto &e : ev) (void)hipEventDestroy(e);
    std::sort(ms.begin(), ms.end());
    return ms[reps / 2];
}

extern "C" int dh_autotune_proton_wet(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                      const float *d_rlsp, int32_t reps, int32_t *best_raytrace,
                                      int32_t *best_smooth, float *ms_raytrace, float *ms_smooth, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(grid && beam && d_rlsp, "null argument");
        dh_device_enter(gpu_id);
        hipStream_t s = (hipStream_t)stream;
        const dh_geom g = dh_proton_geom(grid, beam);
        const size_t n = (size_t)g.nx * g.ny * g.nz;
        const dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
        float4 *vpad = (float4 *)dh_scratch(gpu_id, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
        float *ref = (float *)dh_scratch(gpu_id, DH_SLOT_TMP0, sizeof(float) * n);
        float *out = (float *)dh_scratch(gpu_id, DH_SLOT_TMP1, sizeof(float) * n);
        float *sm_ref = (float *)dh_scratch(gpu_id, DH_SLOT_OUT0, sizeof(float) * n);
        dh_launch_pad_quad(d_rlsp, vpad, g.nx, g.ny, g.nz, 0.0f, s);
        const std::string key = dh_tune_key_geom(g);
        int best = 0;
        float best_ms = 0.0f;
        for (int c = 0; c < DH_TILE_CHOICES; c++) {
            float *o = c == 0 ? ref : out;
            const float ms = dh_time_median([&] { dh_launch_raytrace(g, vpad, o, s, c); }, reps, s);
            if (ms_raytrace) ms_raytrace[c] = ms;
            DH_REQUIRE(c == 0 || dh_count_diff(ref, out, (int64_t)n, s) == 0,
                       "ray-trace variants disagree (tile choice changed the result)");
            if (c == 0 || ms < best_ms) { best = c; best_ms = ms; }
        }
        dh_tune_store("raytrace", key, best, best_ms);
        if (best_raytrace) *best_raytrace = best;
        best = 0;
        for (int c = 0; c < DH_TILE_CHOICES; c++) {
            float *o = c == 0 ? sm_ref : out;
            const float ms = dh_time_median([&] { dh_launch_smooth(g, ref, o, s, c); }, reps, s);
            if (ms_smooth) ms_smooth[c] = ms;
            DH_REQUIRE(c == 0 || dh_count_diff(sm_ref, out, (int64_t)n, s) == 0,
                       "smoothing variants disagree (tile choice changed the result)");
            if (c == 0 || ms < best_ms) { best = c; best_ms = ms; }
        }
        dh_tune_store("smooth", key, best, best_ms);
        if (best_smooth) *best_smooth = best;
    });
}
