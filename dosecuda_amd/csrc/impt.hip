// impt.hip -- proton path on gfx950: volume padding, radiological-depth ray
// trace, lateral WET smoothing and the double-Gaussian pencil-beam sum.
//
// Layout: every volume is the caller's C-order (nz, ny, nx) float32 array;
// consecutive lanes of a wave64 own consecutive x voxels, so every gather of
// a wave touches a few adjacent cache lines.  Sampled volumes are copied once
// into a padded layout (dh_pad) whose border cells hold the texture border
// value, which makes trilinear sampling branch-free.
#include "dh_runtime.h"

namespace {

constexpr int kBlock = 256;

// Quad-packed padded volume (1 voxel of border below, 2 above on each axis):
// element (x, y, z) of the padded grid holds, of plane z, the padded values
// v(x, y) and v(x, y + 1), each followed by its forward x difference
// v(x + 1, .) - v(x, .); outside the volume the value is the border.  Elements
// of the last padded row / column are never read by the samplers.
__global__ void __launch_bounds__(kBlock)
k_pad_quad(const float *__restrict__ src, float4 *__restrict__ dst, int nx, int ny, int nz,
           dh_pad p, float border)
{
    const uint32_t n = (uint32_t)(p.pxy * p.pz);
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const int i = (int)(t % (uint32_t)p.px) - 1;
        const uint32_t r = t / (uint32_t)p.px;
        const int j = (int)(r % (uint32_t)p.py) - 1;
        const int k = (int)(r / (uint32_t)p.py) - 1;
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const int ii = i + (c & 1), jj = j + (c >> 1);
            v[c] = (ii >= 0 && ii < nx && jj >= 0 && jj < ny && k >= 0 && k < nz)
                       ? src[(uint32_t)ii + (uint32_t)nx * ((uint32_t)jj + (uint32_t)ny * (uint32_t)k)]
                       : border;
        }
        dst[t] = make_float4(v[0], v[1] - v[0], v[2], v[3] - v[2]);
    }
}

// Radiological depth to the source (CudaClasses.cu:180-228): samples every
// 1 mm from the voxel centre towards the source until the sample position
// leaves the volume; WET = -0.05 + 0.1 * sum(max(rho, 0)) cm.  The sample
// lattice and exit test are the reference's FP32 arithmetic; the running sum
// is FP64 so long heterogeneous rays do not lose low-order bits.
// Number of leading samples L = 0, 1, ... that are provably inside the volume,
// from the ray/box exit distance evaluated in FP32, minus a margin per axis
// of (2e-3 + 8e-6 n) voxel / |du|.  That margin bounds every rounding here
// and in the sample's own texel coordinate: the texel coordinates are below
// n + 1 (FP32 error < 2 ulp of n, i.e. < 2.4e-7 n voxel), and the quotient
// (lim - t0) / du carries a relative error below 4 ulp, i.e. below
// 5e-7 * La < 5e-7 * n / |du| voxel-steps -- both far below the margin.
// A short bulk only moves samples to the exactly tested tail.
__device__ static inline int safe_samples(const dh_geom &g, float x, float y, float z,
                                          float ux, float uy, float uz, float step)
{
    const float p[3] = {x, y, z}, u[3] = {ux, uy, uz};
    const int n[3] = {g.nx, g.ny, g.nz};
    float Lmin = 1e30f;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float t0 = fmaf(g.inv_s, p[a] + g.iso[a], 0.5f);
        const float du = g.inv_s * u[a];
        if (du == 0.0f) continue;
        const float lim = du > 0.0f ? (float)n[a] : 0.0f;
        const float La = (lim - t0) / du - (2e-3f + 8e-6f * (float)n[a]) / fabsf(du);
        Lmin = fminf(Lmin, La);
    }
    if (!(Lmin >= 0.0f)) return 0;
    return (int)fminf(floorf(Lmin / step) + 1.0f, 1.0e7f);
}

// True when the voxel at image point (x, y, z) lies outside the head-frame box.
__device__ static inline bool dh_outside(const dh_geom &g, const dh_hbox &box, float x, float y, float z)
{
    if (!box.on) return false;
    float hx, hy, hz;
    dh_img_to_head(g, x, y, z, hx, hy, hz);
    return !(hx >= box.x0 && hx <= box.x1 && hy >= box.y0 && hy <= box.y1);
}

// step: sample spacing in mm.  1 is the reference's ray trace; the checked
// surrogate path traces with one sample per voxel (step = spacing), whose
// Riemann sum is WET = -0.05 step + 0.1 step * sum (cm), off the parity path.
template <bool Off32>
__global__ void __launch_bounds__(kBlock)
k_raytrace(dh_geom g, dh_pad p, dh_tile tile, dh_hbox box, float step, const float4 *__restrict__ vol,
           float *__restrict__ wet)
{
    const float4 *vol_hi = vol + p.pxy;
    int i, j, k;
    if (!dh_tile_voxel(tile, dh_xcd_block(blockIdx.x, gridDim.x), threadIdx.x, g.nx, g.ny, g.nz, i, j, k)) return;
    const uint32_t v = (uint32_t)i + (uint32_t)g.nx * ((uint32_t)j + (uint32_t)g.ny * (uint32_t)k);
    float x, y, z, ux, uy, uz;
    dh_vox_xyz(g, i, j, k, x, y, z);
    if (dh_outside(g, box, x, y, z)) return;
    dh_uvec(g, x, y, z, ux, uy, uz);
    const float fnx = (float)g.nx, fny = (float)g.ny, fnz = (float)g.nz;
    double acc = 0.0;
    // Bulk: samples known to be inside need no exit test; their texel
    // coordinate is u = u0 + L * du (continuous in L, no decision rides on it).
    const int nsafe = safe_samples(g, x, y, z, ux, uy, uz, step);
    const float u0x = fmaf(g.inv_s, x + g.iso[0], 0.5f) - 0.5f;
    const float u0y = fmaf(g.inv_s, y + g.iso[1], 0.5f) - 0.5f;
    const float u0z = fmaf(g.inv_s, z + g.iso[2], 0.5f) - 0.5f;
    const float dux = ux * g.inv_s, duy = uy * g.inv_s, duz = uz * g.inv_s;
    int m = 0;
    for (; m + 4 <= nsafe; m += 4) {
        float part = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float L = (float)(m + q) * step;
            part += fmaxf(dh_trilinear_uq<Off32>(vol, vol_hi, p, fmaf(dux, L, u0x), fmaf(duy, L, u0y),
                                                 fmaf(duz, L, u0z)), 0.0f);
        }
        acc += (double)part;
    }
    // Tail: the reference's exact FP32 sample positions and exit test, four
    // at a time.  Inside-ness is monotone along a ray leaving a box, so a
    // sample counts iff every earlier one was inside; loads past the exit read
    // clamped, in-range addresses and are discarded.
    int mt = m;
    for (;;) {
        float d[4];
        bool in[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float Lq = (float)(mt + q) * step;
            const float qx = fmaf(ux, Lq, x), qy = fmaf(uy, Lq, y), qz = fmaf(uz, Lq, z);
            const float tx = fmaf(g.inv_s, qx + g.iso[0], 0.5f);
            const float ty = fmaf(g.inv_s, qy + g.iso[1], 0.5f);
            const float tz = fmaf(g.inv_s, qz + g.iso[2], 0.5f);
            in[q] = tx >= 0.0f && tx < fnx && ty >= 0.0f && ty < fny && tz >= 0.0f && tz < fnz;
            d[q] = dh_trilinear_q<Off32>(vol, vol_hi, p, g.nx, g.ny, g.nz, tx, ty, tz);
        }
        float part = 0.0f;
        bool all = true;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            all = all && in[q];
            part += all ? fmaxf(d[q], 0.0f) : 0.0f;
        }
        acc += (double)part;
        if (!all) break;
        mt += 4;
    }
    wet[v] = (float)(-0.05 * (double)step + 0.1 * (double)step * acc);
}

struct smooth_dirs {
    float c[6], s[6];
    float tx[6], ty[6], tz[6];   // image-index step per mm along each direction
};

// Lateral smoothing (IMPTClasses.cu:139-202): mean of the voxel's WET and the
// nearest-voxel WETs at 1..9 mm along six beam's-eye-view directions, taking
// only radii below 10 * WET(cm).  Neighbour indices come from the reference's
// FP32 head->image rotation and roundf().
__global__ void __launch_bounds__(kBlock)
k_smooth(dh_geom g, smooth_dirs dirs, dh_tile tile, dh_hbox box, const float *__restrict__ wet_in,
         float *__restrict__ wet_out)
{
    int i, j, k;
    if (!dh_tile_voxel(tile, dh_xcd_block(blockIdx.x, gridDim.x), threadIdx.x, g.nx, g.ny, g.nz, i, j, k)) return;
    const uint32_t v = (uint32_t)i + (uint32_t)g.nx * ((uint32_t)j + (uint32_t)g.ny * (uint32_t)k);
    float x, y, z, hx, hy, hz;
    dh_vox_xyz(g, i, j, k, x, y, z);
    dh_img_to_head(g, x, y, z, hx, hy, hz);
    if (box.on && !(hx >= box.x0 && hx <= box.x1 && hy >= box.y0 && hy <= box.y1)) {
        wet_out[v] = DH_WET_OUTSIDE;
        return;
    }
    const float c = wet_in[v];
    // Neumaier-compensated FP32 sum of at most 55 values: the rounded sum is
    // within one ulp of the exact one before the single division
    float sum = c, comp = 0.0f;
    int n = 1;
    const float lim = c * 10.0f;
    // The neighbour's continuous image index is linear in the radius: from
    // this voxel's index t0 in steps of dirs.t per mm it is known to ~1e-4
    // voxel.  Where every coordinate is more than 2e-3 from a rounding tie
    // that fixes the rounded index; lanes near a tie redo the reference's
    // FP32 sequence (rotation, division by the spacing, roundf).
    float q0x, q0y, q0z;
    dh_head_to_img(g, hx, hy, hz, q0x, q0y, q0z);
    const float t0x = (q0x + g.iso[0]) * g.inv_s, t0y = (q0y + g.iso[1]) * g.inv_s;
    const float t0z = (q0z + g.iso[2]) * g.inv_s;
    for (int a = 0; a < 6; a++) {
        for (float dr = 1.0f; dr < lim && dr < 10.0f; dr += 1.0f) {
            const float tx = fmaf(dr, dirs.tx[a], t0x), ty = fmaf(dr, dirs.ty[a], t0y);
            const float tz = fmaf(dr, dirs.tz[a], t0z);
            const float fx = tx - floorf(tx), fy = ty - floorf(ty), fz = tz - floorf(tz);
            const bool tie = fabsf(fx - 0.5f) < 2e-3f || fabsf(fy - 0.5f) < 2e-3f || fabsf(fz - 0.5f) < 2e-3f;
            unsigned ii, jj, kk;
            if (!tie) {
                ii = (unsigned)(int)floorf(tx + 0.5f);
                jj = (unsigned)(int)floorf(ty + 0.5f);
                kk = (unsigned)(int)floorf(tz + 0.5f);
            } else {
                float qx, qy, qz;
                dh_head_to_img(g, hx + dr * dirs.c[a], hy + dr * dirs.s[a], hz, qx, qy, qz);
                ii = (unsigned)(int)roundf((qx + g.iso[0]) / g.s);
                jj = (unsigned)(int)roundf((qy + g.iso[1]) / g.s);
                kk = (unsigned)(int)roundf((qz + g.iso[2]) / g.s);
            }
            if (ii < (unsigned)g.nx && jj < (unsigned)g.ny && kk < (unsigned)g.nz) {
                dh_neumaier(sum, comp, wet_in[ii + (unsigned)g.nx * (jj + (unsigned)g.ny * kk)]);
                n++;
            }
        }
    }
    wet_out[v] = (sum + comp) / (float)n;
}

// Pencil-beam sum (IMPTClasses.cu:204-271).  One lane owns one voxel and walks
// every layer in a fixed order, so there are no atomics and the result is
// bit-reproducible.  Per layer the workgroup stages the layer's IDD/sigma
// table in LDS; per (voxel, layer) the range cut, LUT interpolation, sigma_air,
// nuclear halo and the two Gaussian normalisations are evaluated once; the
// spot loop then reads wave-uniform spot records through the scalar cache.
// With `Fold`, the normalisation is folded into the exponent,
// P*exp(a d^2) = 2^(a' d^2 + log2 P), which requires P, H >= 0 (the host checks
// that every IDD sample of the plan is non-negative).
template <bool Fold, bool Lds>
__global__ void __launch_bounds__(kBlock)
k_pencil(dh_geom g, float vsadx, float vsady, int n_layers, dh_tile tile, const dh_layer_k *__restrict__ layers,
         const float4 *__restrict__ luts, const float *__restrict__ spots,
         const float *__restrict__ wet, float *__restrict__ dose)
{
    __shared__ float4 s_lut[Lds ? DH_LUT_LENGTH : 1];
    int i, j, k;
    const bool valid = dh_tile_voxel(tile, dh_xcd_block(blockIdx.x, gridDim.x), threadIdx.x, g.nx, g.ny, g.nz,
                                     i, j, k);
    if (!valid) { i = g.nx - 1; j = g.ny - 1; k = g.nz - 1; }
    const uint32_t vv = (uint32_t)i + (uint32_t)g.nx * ((uint32_t)j + (uint32_t)g.ny * (uint32_t)k);
    const uint32_t v = vv;
    float x, y, z, hx, hy, hz;
    dh_vox_xyz(g, i, j, k, x, y, z);
    dh_img_to_head(g, x, y, z, hx, hy, hz);
    const float dxs = g.src[0] - x, dys = g.src[1] - y, dzs = g.src[2] - z;
    const float dist = sqrtf(dxs * dxs + dys * dys + dzs * dzs);   // norm3df
    const float w = wet[vv] * 10.0f;                                // cm -> mm
    // The spot axis through (sx, sy, 0) with direction (sx/VSADX, sy/VSADY, -1)
    // crosses the voxel's BEV depth hz at (sx*cx, sy*cy, hz); the squared
    // distance to the axis is |u|^2 - (u.t)^2 for the in-plane offset
    // u = (hx - sx*cx, hy - sy*cy, 0) and the unit axis t: the same quantity as
    // caxDistance (IMPTClasses.cu:123-133) without its hz^2 cancellation.
    const float cx = 1.0f - hz / vsadx, cy = 1.0f - hz / vsady;
    float tot = 0.0f, comp = 0.0f;

    for (int l = 0; l < n_layers; l++) {
        const dh_layer_k L = layers[l];
        const float4 *lut = luts + (size_t)l * DH_LUT_LENGTH;
        if (Lds) {
            __syncthreads();
            for (int q = threadIdx.x; q < DH_LUT_LENGTH; q += kBlock) s_lut[q] = lut[q];
            __syncthreads();
            lut = s_lut;
        }
        const bool act = valid && !(w > L.cut);
        if (!dh_wave_any(act)) continue;
        if (!act) continue;

        // LUT, sigma_air, halo and normalisations (IMPTClasses.cu:69-121,222-264)
        const dh_pb_factors fc = dh_pencil_factors(lut, w, dist, L.r80, L.A, L.B, L.C);
        const float P = fc.P, H = fc.H, ap = fc.ap, ah = fc.ah;
        const float lp = Fold ? log2f(P) : P;
        const float lh = Fold ? log2f(H) : H;

        const float *sp = spots + (size_t)8 * L.start;
        for (int s0 = 0; s0 < L.n; s0 += 16) {
            const int e = min(16, L.n - s0);
            float part = 0.0f;
            for (int u = 0; u < e; u++) {
                const float *rec = sp + 8 * (s0 + u);
                const float sx = rec[0], sy = rec[1], tx = rec[2], ty = rec[3], mu = rec[4];
                const float ux = fmaf(-sx, cx, hx), uy = fmaf(-sy, cy, hy);
                const float r2 = fmaf(ux, ux, uy * uy);
                const float dt = fmaf(tx, ux, ty * uy);
                const float d2 = fmaf(-dt, dt, r2);
                float val;
                if (Fold) {
                    val = __builtin_amdgcn_exp2f(fmaf(ap, d2, lp)) + __builtin_amdgcn_exp2f(fmaf(ah, d2, lh));
                } else {
                    val = fmaf(lp, __builtin_amdgcn_exp2f(ap * d2), lh * __builtin_amdgcn_exp2f(ah * d2));
                }
                part = fmaf(mu, val, part);
            }
            dh_neumaier(tot, comp, part);
        }
    }
    if (valid) dose[v] = tot + comp;
}

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

void dh_launch_pad_quad(const float *src, float4 *dst, int nx, int ny, int nz, float border, hipStream_t s)
{
    dh_pad p = dh_make_pad(nx, ny, nz);
    size_t n = (size_t)p.pxy * p.pz;
    unsigned grid = (unsigned)std::min<size_t>(grid_for(n), 4096);
    dh_timed t("pad", s);
    hipLaunchKernelGGL(k_pad_quad, dim3(grid), dim3(kBlock), 0, s, src, dst, nx, ny, nz, p, border);
    DH_HIP(hipGetLastError());
}

void dh_launch_raytrace(const dh_geom &g, const float4 *vol_pad, float *wet, hipStream_t s, int choice,
                        const dh_hbox *box, float step)
{
    DH_REQUIRE(step >= 1.0f && step <= 16.0f, "ray-trace step must be 1..16 mm");
    const dh_hbox b = box ? *box : dh_hbox{0.0f, 0.0f, 0.0f, 0.0f, 0};
    dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
    if (choice < 0) choice = dh_tune_lookup("raytrace", dh_tune_key_geom(g), 0);
    unsigned grid = 0;
    const dh_tile tile = dh_make_tile(choice, g.nx, g.ny, g.nz, &grid);
    dh_timed t("raytrace", s);
    // 32-bit offsets for quad volumes under 2 GiB whose planes fit 24-bit products
    if ((size_t)16 * p.pxy * p.pz < ((size_t)1 << 31) && p.pxy < ((int64_t)1 << 23))
        hipLaunchKernelGGL(k_raytrace<true>, dim3(grid), dim3(kBlock), 0, s, g, p, tile, b, step, vol_pad, wet);
    else
        hipLaunchKernelGGL(k_raytrace<false>, dim3(grid), dim3(kBlock), 0, s, g, p, tile, b, step, vol_pad, wet);
    DH_HIP(hipGetLastError());
}

void dh_launch_smooth(const dh_geom &g, const float *wet_raw, float *wet_out, hipStream_t s, int choice,
                      const dh_hbox *box)
{
    const dh_hbox b = box ? *box : dh_hbox{0.0f, 0.0f, 0.0f, 0.0f, 0};
    smooth_dirs d;
    for (int a = 0; a < 6; a++) {           // IMPTClasses.cu:170
        float ang = (float)a * DH_PI_F / 3.0f;
        d.s[a] = sinf(ang);
        d.c[a] = cosf(ang);
        // the head->image rotation is linear: image-index step per mm
        float x, y, z;
        dh_head_to_img(g, d.c[a], d.s[a], 0.0f, x, y, z);
        d.tx[a] = x * g.inv_s;
        d.ty[a] = y * g.inv_s;
        d.tz[a] = z * g.inv_s;
    }
    if (choice < 0) choice = dh_tune_lookup("smooth", dh_tune_key_geom(g), 0);
    unsigned grid = 0;
    const dh_tile tile = dh_make_tile(choice, g.nx, g.ny, g.nz, &grid);
    dh_timed t("smooth", s);
    hipLaunchKernelGGL(k_smooth, dim3(grid), dim3(kBlock), 0, s, g, d, tile, b, wet_raw, wet_out);
    DH_HIP(hipGetLastError());
}

void dh_launch_pencil(const dh_geom &g, float vsadx, float vsady, int n_layers, const dh_layer_k *d_layers,
                      const float4 *d_luts, const float *d_spots, bool fold, const float *wet,
                      float *dose, hipStream_t s, int choice)
{
    if (choice < 0) choice = dh_tune_lookup("pencil", dh_tune_key_geom(g), 0);
    DH_REQUIRE(choice >= 0 && choice < 2 * DH_TILE_CHOICES, "pencil choice out of range");
    const bool lds = choice < DH_TILE_CHOICES;
    unsigned grid = 0;
    const dh_tile tile = dh_make_tile(choice % DH_TILE_CHOICES, g.nx, g.ny, g.nz, &grid);
    dh_timed t("pencil", s);
    if (fold && lds)
        hipLaunchKernelGGL((k_pencil<true, true>), dim3(grid), dim3(kBlock), 0, s, g, vsadx, vsady, n_layers, tile,
                           d_layers, d_luts, d_spots, wet, dose);
    else if (fold)
        hipLaunchKernelGGL((k_pencil<true, false>), dim3(grid), dim3(kBlock), 0, s, g, vsadx, vsady, n_layers, tile,
                           d_layers, d_luts, d_spots, wet, dose);
    else if (lds)
        hipLaunchKernelGGL((k_pencil<false, true>), dim3(grid), dim3(kBlock), 0, s, g, vsadx, vsady, n_layers, tile,
                           d_layers, d_luts, d_spots, wet, dose);
    else
        hipLaunchKernelGGL((k_pencil<false, false>), dim3(grid), dim3(kBlock), 0, s, g, vsadx, vsady, n_layers,
                           tile, d_layers, d_luts, d_spots, wet, dose);
    DH_HIP(hipGetLastError());
}
