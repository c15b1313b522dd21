// dh_common.h -- shared device/host definitions of the MI355X dose engine.
//
// Everything here is compiled with -ffp-contract=off: the FP32 geometry that
// decides which samples exist (ray exit, rounded neighbour indices) is written
// as explicit multiplies/adds, and explicit fmaf() where the reference calls
// fmaf (DoseCUDA/dose_kernels/CudaClasses.cuh:100-104, CudaClasses.cu:201-216).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define DH_PI_F 3.141592654f  // CUDART_PI_F, used by the reference for angles
#define DH_LOG2E_F 1.44269504088896340736f

// Padded volume: one border voxel below and two above on every axis, so that
// trilinear corners of any clamped coordinate u in [-1, n] are in range and
// out-of-image corners read the border value (TextureClasses.cu:22-31).
struct dh_pad {
    int px, py, pz;      // nx + 3, ny + 3, nz + 3
    int64_t pxy;         // px * py
};

static inline dh_pad dh_make_pad(int nx, int ny, int nz)
{
    dh_pad p;
    p.px = nx + 3; p.py = ny + 3; p.pz = nz + 3;
    p.pxy = (int64_t)p.px * p.py;
    return p;
}

// Per-beam geometry, passed by value as a kernel argument (lands in SGPRs).
struct dh_geom {
    int nx, ny, nz;
    float s, inv_s;
    float iso[3];        // iso - origin, mm
    float src[3];        // source position relative to iso, mm
    float singa, cosga, sinta, costa, sinca, cosca;
};

// ---------------------------------------------------------------------------
// FP32 geometry shared by all kernels (host + device).
// ---------------------------------------------------------------------------
__host__ __device__ static inline void dh_vox_xyz(const dh_geom &g, int i, int j, int k,
                                                  float &x, float &y, float &z)
{   // CudaClasses.cuh:106-110
    x = (float)i * g.s - g.iso[0];
    y = (float)j * g.s - g.iso[1];
    z = (float)k * g.s - g.iso[2];
}

// Unit vector towards the source (CudaClasses.cuh:41-52).  The norm is the
// correctly rounded reciprocal square root of the FP32 sum of squares, taken
// through FP64 so host and device agree bit for bit.
__host__ __device__ static inline void dh_uvec(const dh_geom &g, float x, float y, float z,
                                               float &ux, float &uy, float &uz)
{
    float dx = g.src[0] - x, dy = g.src[1] - y, dz = g.src[2] - z;
    float ss = dx * dx + dy * dy + dz * dz;
    float r = (float)(1.0 / sqrt((double)ss));
    ux = dx * r; uy = dy * r; uz = dz * r;
}

// Image -> head with couch, gantry and collimator rotations
// (IMRTClasses.cu:41-79); sinca = 0, cosca = 1 reduces it exactly to the
// proton version (CudaClasses.cu:65-95).
__host__ __device__ static inline void dh_img_to_head(const dh_geom &g, float x, float y, float z,
                                                      float &hx, float &hy, float &hz)
{
    float sx = -g.sinta, cx = g.costa;
    float xt = x * cx + z * sx;
    float yt = y;
    float zt = -x * sx + z * cx;
    sx = -g.singa; cx = g.cosga;
    float xg = xt * cx - yt * sx;
    float yg = xt * sx + yt * cx;
    float zg = zt;
    sx = -g.sinca; cx = g.cosca;
    float xc = xg * cx + zg * sx;
    float yc = yg;
    float zc = -xg * sx + zg * cx;
    hx = -xc; hy = zc; hz = yc;
}

// Head -> image (IMRTClasses.cu:81-118; CudaClasses.cu:97-126).
__host__ __device__ static inline void dh_head_to_img(const dh_geom &g, float hx, float hy, float hz,
                                                      float &x, float &y, float &z)
{
    float xz = -hx, yz = hz, zz = hy;
    float sx = g.sinca, cx = g.cosca;
    float xc = xz * cx + zz * sx;
    float yc = yz;
    float zc = -xz * sx + zz * cx;
    sx = g.singa; cx = g.cosga;
    float xg = xc * cx - yc * sx;
    float yg = xc * sx + yc * cx;
    float zg = zc;
    sx = g.sinta; cx = g.costa;
    x = xg * cx + zg * sx;
    y = yg;
    z = -xg * sx + zg * cx;
}

// Linear index of voxel (i, j, k) in the padded layout.
__host__ __device__ static inline int64_t dh_pad_index(const dh_pad &p, int i, int j, int k)
{
    return (int64_t)(i + 1) + (int64_t)p.px * (j + 1) + p.pxy * (k + 1);
}

#ifdef __HIPCC__
// Trilinear sample of a padded scalar volume at texel coordinate (tx,ty,tz)
// (texel centre of voxel i at t = i + 0.5).  u = t - 0.5 is clamped to
// [-1, n]; outside that range every corner is border anyway.
__device__ static inline float dh_trilinear(const float *__restrict__ vol, const dh_pad &p,
                                            int nx, int ny, int nz,
                                            float tx, float ty, float tz)
{
    float ux = __builtin_amdgcn_fmed3f(tx - 0.5f, -1.0f, (float)nx);
    float uy = __builtin_amdgcn_fmed3f(ty - 0.5f, -1.0f, (float)ny);
    float uz = __builtin_amdgcn_fmed3f(tz - 0.5f, -1.0f, (float)nz);
    float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
    float ax = ux - fx0, ay = uy - fy0, az = uz - fz0;
    int i0 = (int)fx0, j0 = (int)fy0, k0 = (int)fz0;
    const float *b = vol + dh_pad_index(p, i0, j0, k0);
    const float *c = b + p.pxy;
    float v000 = b[0], v100 = b[1], v010 = b[p.px], v110 = b[p.px + 1];
    float v001 = c[0], v101 = c[1], v011 = c[p.px], v111 = c[p.px + 1];
    float a00 = fmaf(ax, v100 - v000, v000);
    float a10 = fmaf(ax, v110 - v010, v010);
    float a01 = fmaf(ax, v101 - v001, v001);
    float a11 = fmaf(ax, v111 - v011, v011);
    float b0 = fmaf(ay, a10 - a00, a00);
    float b1 = fmaf(ay, a11 - a01, a01);
    return fmaf(az, b1 - b0, b0);
}

// XCD-aware workgroup numbering.  The dispatcher hands workgroup b to XCD
// b % 8; this maps b to a logical index so that each XCD receives one
// contiguous range of logical workgroups (and so one compact region of the
// volume, which then stays in that XCD's own L2).  Bijective on [0, nb).
constexpr int DH_NUM_XCD = 8;
__device__ static inline int dh_xcd_block(int b, int nb)
{
    const int x = b % DH_NUM_XCD, q = b / DH_NUM_XCD;
    const int base = nb / DH_NUM_XCD, rem = nb % DH_NUM_XCD;
    return x * base + min(x, rem) + q;
}

// Voxels of a 256-thread workgroup: linear runs along x (tx == 0, the
// voxel index is b * 256 + thread) or a tx x ty x tz brick (product 256),
// bricks numbered x fastest (yfirst = 0) or y fastest (yfirst = 1; then x,
// then z).  The numbering decides which bricks run at the same time: for rays
// along y, y-fastest bricks share their rays' columns, so their working set
// stays within one XCD's L2.  Returns whether the voxel is in the volume.
struct dh_tile {
    int tx, ty, tz, ntx, nty, yfirst;
};

__device__ static inline bool dh_tile_voxel(const dh_tile &t, int b, int tid, int nx, int ny, int nz, int &i,
                                            int &j, int &k)
{
    if (t.tx == 0) {
        const uint32_t v = (uint32_t)b * 256u + (uint32_t)tid;
        i = (int)(v % (uint32_t)nx);
        const uint32_t r = v / (uint32_t)nx;
        j = (int)(r % (uint32_t)ny);
        k = (int)(r / (uint32_t)ny);
        return v < (uint32_t)nx * ny * nz;
    }
    int bx, by, bz;
    if (t.yfirst) {
        by = b % t.nty;
        const int r = b / t.nty;
        bx = r % t.ntx;
        bz = r / t.ntx;
    } else {
        bx = b % t.ntx;
        const int r = b / t.ntx;
        by = r % t.nty;
        bz = r / t.nty;
    }
    i = bx * t.tx + tid % t.tx;
    j = by * t.ty + (tid / t.tx) % t.ty;
    k = bz * t.tz + tid / (t.tx * t.ty);
    return i < nx && j < ny && k < nz;
}

// Ordered compaction by a 256-thread workgroup over K rounds of 256
// consecutive items starting at `base`: item u stands for the value map(u)
// (negative: none), flag(map(u)) is evaluated once per item, and the flagged
// values are appended to `list` in item order with one atomic per workgroup
// (K times fewer than one per round).
template <int K, class Map, class Flag>
__device__ static inline void dh_block_compact_map(uint32_t base, uint32_t n, Map map, Flag flag,
                                                   int *__restrict__ list, int *__restrict__ counter)
{
    __shared__ int s_cnt[K][4];
    __shared__ int s_base;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    uint64_t m[K];
    int val[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        const uint32_t u = base + (uint32_t)q * 256u + threadIdx.x;
        val[q] = u < n ? map(u) : -1;
        const bool f = val[q] >= 0 && flag((uint32_t)val[q]);
        m[q] = __builtin_amdgcn_ballot_w64(f);
        if (lane == 0) s_cnt[q][wave] = __popcll(m[q]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int q = 0; q < K; q++)
            for (int w = 0; w < 4; w++) {
                const int c = s_cnt[q][w];
                s_cnt[q][w] = tot;
                tot += c;
            }
        s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int q = 0; q < K; q++)
        if ((m[q] >> lane) & 1ull) list[s_base + s_cnt[q][wave] + __popcll(m[q] & below)] = val[q];
}

template <int K, class Flag>
__device__ static inline void dh_block_compact(uint32_t base, uint32_t n, Flag flag, int *__restrict__ list,
                                               int *__restrict__ counter)
{
    __shared__ int s_cnt[K][4];
    __shared__ int s_base;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    uint64_t m[K];
#pragma unroll
    for (int q = 0; q < K; q++) {
        const uint32_t v = base + (uint32_t)q * 256u + threadIdx.x;
        const bool f = v < n && flag(v);
        m[q] = __builtin_amdgcn_ballot_w64(f);
        if (lane == 0) s_cnt[q][wave] = __popcll(m[q]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int q = 0; q < K; q++)
            for (int w = 0; w < 4; w++) {
                const int c = s_cnt[q][w];
                s_cnt[q][w] = tot;
                tot += c;
            }
        s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int q = 0; q < K; q++)
        if ((m[q] >> lane) & 1ull)
            list[s_base + s_cnt[q][wave] + __popcll(m[q] & below)] = (int)(base + (uint32_t)q * 256u + threadIdx.x);
}

// Trilinear sample at a texel-space point u = t - 0.5 known to lie in
// [-0.5, n - 0.5) on every axis (no clamping needed).
__device__ static inline float dh_trilinear_u(const float *__restrict__ vol, const dh_pad &p,
                                              float ux, float uy, float uz)
{
    float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
    float ax = ux - fx0, ay = uy - fy0, az = uz - fz0;
    const float *b = vol + dh_pad_index(p, (int)fx0, (int)fy0, (int)fz0);
    const float *c = b + p.pxy;
    float v000 = b[0], v100 = b[1], v010 = b[p.px], v110 = b[p.px + 1];
    float v001 = c[0], v101 = c[1], v011 = c[p.px], v111 = c[p.px + 1];
    float a00 = fmaf(ax, v100 - v000, v000);
    float a10 = fmaf(ax, v110 - v010, v010);
    float a01 = fmaf(ax, v101 - v001, v001);
    float a11 = fmaf(ax, v111 - v011, v011);
    float b0 = fmaf(ay, a10 - a00, a00);
    float b1 = fmaf(ay, a11 - a01, a01);
    return fmaf(az, b1 - b0, b0);
}

// The same two samplers over the quad-packed padded volume: element e holds
// {v(e), v(e + 1) - v(e), v(e + px), v(e + px + 1) - v(e + px)}, so the 8
// corners come in two 16-byte loads (planes k0 and k0 + 1) instead of eight,
// and each x interpolation fmaf(a, v1 - v0, v0) reads the difference it would
// have formed itself: the results are the same as from the plain corners.
__device__ static inline float dh_blend8(float4 lo, float4 hi, float ax, float ay, float az)
{
    float a00 = fmaf(ax, lo.y, lo.x);
    float a10 = fmaf(ax, lo.w, lo.z);
    float a01 = fmaf(ax, hi.y, hi.x);
    float a11 = fmaf(ax, hi.w, hi.z);
    float b0 = fmaf(ay, a10 - a00, a00);
    float b1 = fmaf(ay, a11 - a01, a01);
    return fmaf(az, b1 - b0, b0);
}

// Element index of padded voxel (i + 1, j + 1, k + 1), i, j, k >= -1, in 32
// bits (24-bit products): valid where px * py < 2^23 and the volume has fewer
// than 2^31 elements.
__device__ static inline uint32_t dh_pad_index32(const dh_pad &p, int i, int j, int k)
{
    const int pxy = (int)p.pxy;
    return (uint32_t)(__mul24(k, pxy) + __mul24(j, p.px) + i + (1 + p.px + pxy));
}

// Trilinear sample of the quad-packed volume at u = t - 0.5 (already clamped
// or known inside); q_hi = q + pxy.  Off32 takes 32-bit byte offsets against
// the two uniform plane bases (volumes under 2 GiB, px * py < 2^23).
template <bool Off32>
__device__ static inline float dh_trilinear_uq(const float4 *__restrict__ q, const float4 *__restrict__ q_hi,
                                               const dh_pad &p, float ux, float uy, float uz)
{
    float fx0 = floorf(ux), fy0 = floorf(uy), fz0 = floorf(uz);
    float ax = ux - fx0, ay = uy - fy0, az = uz - fz0;
    if constexpr (Off32) {
        const uint32_t off = dh_pad_index32(p, (int)fx0, (int)fy0, (int)fz0) * 16u;
        return dh_blend8(*(const float4 *)((const char *)q + off), *(const float4 *)((const char *)q_hi + off),
                         ax, ay, az);
    } else {
        const int64_t e = dh_pad_index(p, (int)fx0, (int)fy0, (int)fz0);
        return dh_blend8(q[e], q_hi[e], ax, ay, az);
    }
}

template <bool Off32>
__device__ static inline float dh_trilinear_q(const float4 *__restrict__ q, const float4 *__restrict__ q_hi,
                                              const dh_pad &p, int nx, int ny, int nz, float tx, float ty, float tz)
{
    float ux = __builtin_amdgcn_fmed3f(tx - 0.5f, -1.0f, (float)nx);
    float uy = __builtin_amdgcn_fmed3f(ty - 0.5f, -1.0f, (float)ny);
    float uz = __builtin_amdgcn_fmed3f(tz - 0.5f, -1.0f, (float)nz);
    return dh_trilinear_uq<Off32>(q, q_hi, p, ux, uy, uz);
}

// Per (voxel, layer) factors of the double-Gaussian pencil beam
// (IMPTClasses.cu:69-121,222-264): LUT lookup at WET w (mm) in a 400-entry
// {depth, sigma, idd, -} table, sigma_air at source distance dist, nuclear
// halo, and the two Gaussians' normalisations P, H and exponent scales
// ap, ah (already multiplied by log2 e, so exp(a d^2) = 2^(ap d^2)).
struct dh_pb_factors {
    float P, H, ap, ah, st, sht;
};

__device__ static inline dh_pb_factors dh_pencil_factors(const float4 *lut, float w, float dist,
                                                         float r80, float A, float B, float C)
{
    int lo = 0, hi = 400;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (w <= lut[m].x) hi = m; else lo = m + 1;
    }
    float idd, sms;
    if (lo == 400) { idd = lut[lo - 1].z; sms = lut[lo - 1].y; }
    else if (lo == 0) { idd = lut[0].z; sms = lut[0].y; }
    else {
        const float4 a = lut[lo - 1], b = lut[lo];
        const float f = (w - a.x) / (b.x - a.x);
        idd = fmaf(b.z - a.z, f, a.z);
        sms = fmaf(b.y - a.y, f, a.y);
    }
    const float da = dist - w + 0.7f * r80;                      // sigmaAir
    const float st = fmaf(fmaf(A, da, B), da, C) + sms;
    const float wc = fminf(fmaxf(w, 0.1f), r80 - 0.1f);         // nuclearHalo
    const float t5 = wc - r80 - 5.0f;
    float sh = 2.85f + (0.0014f * r80 * logf(wc + 3.0f)) + (0.06f * wc) - (7.4e-5f * (wc * wc))
             - ((0.22f * r80) / (t5 * t5));
    const float r3 = r80 + 3.0f;
    float hw = 0.052f * logf(1.13f + (wc / (11.2f - (0.023f * r80))))
             + (0.35f * ((0.0017f * (r80 * r80)) - r80) / ((r3 * r3) - (wc * wc)))
             - (1.61e-9f * wc * (r3 * r3));
    sh = fmaxf(sh, 0.0f);
    hw = fminf(fmaxf(hw, 0.0f), 0.9f);
    dh_pb_factors f;
    f.st = st;
    f.sht = hypotf(st, sh);
    f.P = (1.0f - hw) * idd / (2.0f * DH_PI_F * (st * st));
    f.H = hw * idd / (2.0f * DH_PI_F * (f.sht * f.sht));
    f.ap = st != 0.0f ? (-0.5f / (st * st)) * DH_LOG2E_F : -INFINITY;
    f.ah = f.sht != 0.0f ? (-0.5f / (f.sht * f.sht)) * DH_LOG2E_F : -INFINITY;
    return f;
}

// Wave-level helpers (wave64).
__device__ static inline bool dh_wave_any(bool pred)
{
    return __builtin_amdgcn_ballot_w64(pred) != 0ull;
}

// Neumaier-compensated accumulation of x into (s, c).
__device__ static inline void dh_neumaier(float &s, float &c, float x)
{
    float t = s + x;
    if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}
#endif
