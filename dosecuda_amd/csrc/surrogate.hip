// surrogate.hip -- BF16 MFMA kernels of the learned CCCS surrogate (a small
// 3-D convolutional encoder-decoder), plus the gamma index used to judge it.
//
// Tensors are channels-last bf16: element (n, z, y, x, c) at
// ((((n * D + z) * H + y) * W + x) * C + c).  Convolutions are implicit GEMMs
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation: a wave owns a block of
// 32 consecutive output voxels x 32 output channels; the K loop runs over
// taps x 16-channel slices.  In the 32x32x16 operand layout, lane l holds row
// (l mod 32) and the 8 consecutive K elements 8*(l/32) .. 8*(l/32)+7, so with
// channels-last activations and [tap][cout][cin] weights every A and B
// fragment is one 16-byte load.  Padding is "same" (zero outside).
#include "dh_runtime.h"

#include <hip/hip_bf16.h>

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 dh_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float dh_f32x16;

namespace {

__device__ static inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ static inline uint16_t f2bf(float f)
{   // round to nearest even (NaN kept quiet)
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

struct conv_args {
    const uint16_t *x;       // [N][D][H][W][cin] bf16
    const uint16_t *w;       // [taps][cout_pad][cin] bf16
    const float *bias;       // [cout] or null
    uint16_t *y;             // [N][D][H][W][cout] bf16 or null
    float *y32;              // same, fp32, or null
    int N, D, H, W, cin, cout, cout_pad, ksize, act;
    float slope;
};

// Forward convolution, k = 1 or 3, stride 1, "same" zero padding.
__global__ void __launch_bounds__(256)
k_conv3d(conv_args a)
{
    const int lane = __lane_id();
    const int wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
    const int nblk_co = a.cout_pad / 32;
    const int64_t nblk_v = (nvox + 31) / 32;
    if (wave >= nblk_v * nblk_co) return;
    const int cob = wave % nblk_co;
    const int64_t vb = wave / nblk_co;
    const int row = lane & 31, kh = lane >> 5;
    // this lane's A row: voxel vb*32 + row
    const int64_t v = vb * 32 + row;
    const bool vin = v < nvox;
    const int64_t vv = vin ? v : nvox - 1;
    const int x = (int)(vv % a.W);
    int64_t t = vv / a.W;
    const int y = (int)(t % a.H); t /= a.H;
    const int z = (int)(t % a.D);
    const int n = (int)(t / a.D);
    const int taps = a.ksize * a.ksize * a.ksize;
    const int rad = a.ksize / 2;
    dh_f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    const int co = cob * 32 + row;       // this lane's B column
    for (int tp = 0; tp < taps; tp++) {
        const int dz = tp / (a.ksize * a.ksize) - rad;
        const int dy = (tp / a.ksize) % a.ksize - rad;
        const int dx = tp % a.ksize - rad;
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        const bool ok = vin && zz >= 0 && zz < a.D && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
        const int64_t src = ok ? ((((int64_t)n * a.D + zz) * a.H + yy) * a.W + xx) : 0;
        const uint16_t *xp = a.x + src * a.cin;
        const uint16_t *wp = a.w + ((int64_t)tp * a.cout_pad + co) * a.cin;
        for (int c0 = 0; c0 < a.cin; c0 += 16) {
            const int cc = c0 + kh * 8;
            uint4 av = ok ? *(const uint4 *)(xp + cc) : make_uint4(0, 0, 0, 0);
            uint4 bv = *(const uint4 *)(wp + cc);
            dh_bf16x8 A, B;
            __builtin_memcpy(&A, &av, 16);
            __builtin_memcpy(&B, &bv, 16);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc, 0, 0, 0);
        }
    }
    // D layout: lane l, register r -> row 8*(r/4) + 4*(l/32) + (r%4), column l%32
    const int col = cob * 32 + row;
    const float b = (a.bias && col < a.cout) ? a.bias[col] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = 8 * (r >> 2) + 4 * kh + (r & 3);
        const int64_t ov = vb * 32 + m;
        if (ov >= nvox || col >= a.cout) continue;
        float o = acc[r] + b;
        if (a.act) o = o > 0.0f ? o : o * a.slope;
        if (a.y) a.y[ov * a.cout + col] = f2bf(o);
        if (a.y32) a.y32[ov * a.cout + col] = o;
    }
}

struct wgrad_args {
    const uint16_t *x;       // [N][D][H][W][cin] bf16 (forward input)
    const uint16_t *dy;      // [N][D][H][W][cout] bf16 (output gradient)
    float *dw;               // [taps][cin][cout] fp32, accumulated with atomics
    int N, D, H, W, cin, cout, ksize;
    int vchunk;              // voxels per work item (multiple of 16)
};

// Weight gradient dW[t][ci][co] = sum_v x[v + off_t][ci] * dy[v][co]:
// M = cin, N = cout, K = voxels.  A work item is (tap, 32x32 (ci, co) block,
// voxel chunk); partial tiles are added to dW with FP32 atomics.
__global__ void __launch_bounds__(256)
k_conv3d_wgrad(wgrad_args a)
{
    const int lane = __lane_id();
    const int64_t wave = (blockIdx.x * 256 + threadIdx.x) >> 6;
    const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
    const int taps = a.ksize * a.ksize * a.ksize;
    const int nci = (a.cin + 31) / 32, nco = (a.cout + 31) / 32;
    const int64_t nchunk = (nvox + a.vchunk - 1) / a.vchunk;
    const int64_t total = (int64_t)taps * nci * nco * nchunk;
    if (wave >= total) return;
    int64_t q = wave;
    const int64_t ch = q % nchunk; q /= nchunk;
    const int cob = (int)(q % nco); q /= nco;
    const int cib = (int)(q % nci); q /= nci;
    const int tp = (int)q;
    const int rad = a.ksize / 2;
    const int dz = tp / (a.ksize * a.ksize) - rad;
    const int dyo = (tp / a.ksize) % a.ksize - rad;
    const int dx = tp % a.ksize - rad;
    const int row = lane & 31, kh = lane >> 5;
    const int ci = cib * 32 + row, co = cob * 32 + row;
    dh_f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    const int64_t v0 = ch * a.vchunk, v1 = min(nvox, v0 + a.vchunk);
    for (int64_t vk = v0; vk < v1; vk += 16) {
        uint16_t ae[8], be[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t v = vk + kh * 8 + e;
            uint16_t av = 0, bv = 0;
            if (v < v1) {
                const int x = (int)(v % a.W);
                int64_t t = v / a.W;
                const int y = (int)(t % a.H); t /= a.H;
                const int z = (int)(t % a.D);
                const int n = (int)(t / a.D);
                const int zz = z + dz, yy = y + dyo, xx = x + dx;
                if (ci < a.cin && zz >= 0 && zz < a.D && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                    av = a.x[((((int64_t)n * a.D + zz) * a.H + yy) * a.W + xx) * a.cin + ci];
                if (co < a.cout) bv = a.dy[v * a.cout + co];
            }
            ae[e] = av;
            be[e] = bv;
        }
        dh_bf16x8 A, B;
        __builtin_memcpy(&A, ae, 16);
        __builtin_memcpy(&B, be, 16);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = 8 * (r >> 2) + 4 * kh + (r & 3);
        const int cim = cib * 32 + m, con = cob * 32 + row;
        if (cim < a.cin && con < a.cout)
            atomicAdd(&a.dw[((int64_t)tp * a.cin + cim) * a.cout + con], acc[r]);
    }
}

// fp32 master weights [taps][cin][cout] -> bf16 forward layout
// [taps][cout_pad][cin] (transpose=0) or the input-gradient layout
// [taps][cin_pad][cout] with the tap order reversed (transpose=1).
__global__ void k_prep_weights(const float *__restrict__ w, int taps, int cin, int cout, int mode,
                               int rows_pad, uint16_t *__restrict__ out)
{
    const int64_t n = (int64_t)taps * rows_pad * (mode ? cout : cin);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (mode == 0) {            // out[t][co][ci] = w[t][ci][co]
            const int ci = (int)(i % cin);
            const int64_t r = i / cin;
            const int co = (int)(r % rows_pad);
            const int t = (int)(r / rows_pad);
            out[i] = co < cout ? f2bf(w[((int64_t)t * cin + ci) * cout + co]) : (uint16_t)0;
        } else {                    // out[t'][ci][co] = w[taps-1-t'][ci][co]
            const int co = (int)(i % cout);
            const int64_t r = i / cout;
            const int ci = (int)(r % rows_pad);
            const int t = (int)(r / rows_pad);
            out[i] = ci < cin ? f2bf(w[((int64_t)(taps - 1 - t) * cin + ci) * cout + co]) : (uint16_t)0;
        }
    }
}

// Gamma index (Low et al. 1998), global normalisation: for each evaluated
// voxel with reference dose above the low-dose cut, the minimum over
// reference voxels within the search radius of
// sqrt(|dr|^2 / dta^2 + (D_eval - D_ref)^2 / (dd * D_norm)^2), evaluated on the
// voxel grid.  Voxels below the cut get -1 (not counted).
__global__ void __launch_bounds__(256)
k_gamma(const float *__restrict__ eval, const float *__restrict__ ref, int nx, int ny, int nz,
        float spacing, float dta, float dd_abs, float cut, int rad, float *__restrict__ out)
{
    const int64_t nvox = (int64_t)nx * ny * nz;
    const int64_t v = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (v >= nvox) return;
    const int i = (int)(v % nx);
    const int64_t r = v / nx;
    const int j = (int)(r % ny), k = (int)(r / ny);
    const float de = eval[v];
    if (!(ref[v] >= cut)) { out[v] = -1.0f; return; }
    const float ia = 1.0f / (dta * dta), ib = 1.0f / (dd_abs * dd_abs);
    float best = 1e30f;
    for (int dk = -rad; dk <= rad; dk++) {
        const int kk = k + dk;
        if (kk < 0 || kk >= nz) continue;
        for (int dj = -rad; dj <= rad; dj++) {
            const int jj = j + dj;
            if (jj < 0 || jj >= ny) continue;
            for (int di = -rad; di <= rad; di++) {
                const int ii = i + di;
                if (ii < 0 || ii >= nx) continue;
                const float d2 = (float)(di * di + dj * dj + dk * dk) * spacing * spacing;
                const float dr = ref[ii + (int64_t)nx * (jj + (int64_t)ny * kk)] - de;
                best = fminf(best, d2 * ia + dr * dr * ib);
            }
        }
    }
    out[v] = sqrtf(best);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int dh_nn_conv3d_dev(const void *x, int32_t n, int32_t d, int32_t h, int32_t w,
                                int32_t cin, const void *wt, const float *bias, int32_t cout,
                                int32_t cout_pad, int32_t ksize, int32_t act, float slope, void *y,
                                float *y32, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(x && wt && (y || y32), "null argument");
        DH_REQUIRE(ksize == 1 || ksize == 3, "kernel size must be 1 or 3");
        DH_REQUIRE(cin > 0 && cin % 16 == 0, "input channels must be a positive multiple of 16");
        DH_REQUIRE(cout > 0 && cout_pad >= cout && cout_pad % 32 == 0, "bad output channel padding");
        DH_REQUIRE(n > 0 && d > 0 && h > 0 && w > 0, "bad shape");
        conv_args a;
        a.x = (const uint16_t *)x; a.w = (const uint16_t *)wt; a.bias = bias;
        a.y = (uint16_t *)y; a.y32 = y32;
        a.N = n; a.D = d; a.H = h; a.W = w; a.cin = cin; a.cout = cout; a.cout_pad = cout_pad;
        a.ksize = ksize; a.act = act; a.slope = slope;
        const int64_t nvox = (int64_t)n * d * h * w;
        const int64_t waves = ((nvox + 31) / 32) * (cout_pad / 32);
        const unsigned grid = (unsigned)((waves + 3) / 4);
        dh_timed t("nn_conv3d", (hipStream_t)stream);
        hipLaunchKernelGGL(k_conv3d, dim3(grid), dim3(256), 0, (hipStream_t)stream, a);
        DH_HIP(hipGetLastError());
    });
}

extern "C" int dh_nn_conv3d_wgrad_dev(const void *x, const void *dy, int32_t n, int32_t d,
                                      int32_t h, int32_t w, int32_t cin, int32_t cout,
                                      int32_t ksize, float *dw, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(x && dy && dw, "null argument");
        DH_REQUIRE(ksize == 1 || ksize == 3, "kernel size must be 1 or 3");
        DH_REQUIRE(cin > 0 && cout > 0 && n > 0 && d > 0 && h > 0 && w > 0, "bad shape");
        wgrad_args a;
        a.x = (const uint16_t *)x; a.dy = (const uint16_t *)dy; a.dw = dw;
        a.N = n; a.D = d; a.H = h; a.W = w; a.cin = cin; a.cout = cout; a.ksize = ksize;
        const int64_t nvox = (int64_t)n * d * h * w;
        a.vchunk = 2048;
        const int taps = ksize * ksize * ksize;
        const int64_t waves = (int64_t)taps * ((cin + 31) / 32) * ((cout + 31) / 32) *
                              ((nvox + a.vchunk - 1) / a.vchunk);
        dh_timed t("nn_conv3d_wgrad", (hipStream_t)stream);
        hipLaunchKernelGGL(k_conv3d_wgrad, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0,
                           (hipStream_t)stream, a);
        DH_HIP(hipGetLastError());
    });
}

extern "C" int dh_nn_prep_weights_dev(const float *w, int32_t taps, int32_t cin, int32_t cout,
                                      int32_t mode, int32_t rows_pad, void *out, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(w && out && taps > 0 && cin > 0 && cout > 0, "bad argument");
        DH_REQUIRE(mode == 0 ? rows_pad >= cout : rows_pad >= cin, "bad padding");
        hipLaunchKernelGGL(k_prep_weights, dim3(256), dim3(256), 0, (hipStream_t)stream, w, taps, cin,
                           cout, mode, rows_pad, (uint16_t *)out);
        DH_HIP(hipGetLastError());
    });
}

extern "C" int dh_gamma_dev(const float *eval, const float *ref, int32_t nz, int32_t ny, int32_t nx,
                            float spacing, float dta_mm, float dd_abs, float cut, int32_t radius_vox,
                            float *out, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(eval && ref && out && nx > 0 && ny > 0 && nz > 0, "bad argument");
        DH_REQUIRE(radius_vox >= 0 && radius_vox <= 16, "search radius must be 0..16 voxels");
        const int64_t n = (int64_t)nx * ny * nz;
        dh_timed t("gamma", (hipStream_t)stream);
        hipLaunchKernelGGL(k_gamma, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                           eval, ref, nx, ny, nz, spacing, dta_mm, dd_abs, cut, radius_vox, out);
        DH_HIP(hipGetLastError());
    });
}
