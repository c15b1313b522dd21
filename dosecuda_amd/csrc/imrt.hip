// imrt.hip -- photon path on gfx950: TERMA / electron contamination and the
// 72-cone collapsed-cone superposition.
//
// The CCCS samples TERMA and density at the same point, so both are packed
// into one padded volume {TERMA (border 0), density (border 0.0012)} read as
// 16-byte gathers: one per trilinear x-row serves both fields.  Voxels with
// TERMA <= 0.01 are written as zero by a compaction pass that also builds the
// list of voxels the cone kernel visits (wave ballots, one atomic per
// workgroup), so the cone kernel's waves are full of working lanes.
#include "dh_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

constexpr int kBlock = 256;
constexpr int kCompactRounds = 8;     // voxels per thread of the compaction

inline unsigned grid_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned grid_compact(size_t n) { return (unsigned)((n + kBlock * kCompactRounds - 1) / (kBlock * kCompactRounds)); }

// erff(x) is exactly +1 for x >= 4.5 and -1 for x <= -4.5 in FP32
// (erfc(4.5) = 2e-10 is far below half an ulp of 1)
constexpr float kErfSaturated = 4.5f;

__device__ static inline bool erf_pair_zero(float a, float b)
{
    return (a >= kErfSaturated && b >= kErfSaturated) || (a <= -kErfSaturated && b <= -kErfSaturated);
}

// Projected, scaled y edge of leaf-pair edge ey for a voxel: the same FP32
// sequence as the transmission sum, monotone in ey when dv > 0.
__device__ static inline float pair_y(float ey, float hy, float dv, float inv)
{
    return (((ey - hy) * dv) + hy) * inv;
}

// Source-occlusion head transmission (IMRTClasses.cu:120-150) over the open
// leaf pairs; the scaled leaf edges {xL, xR, yB, yT} are precomputed per pair.
// A pair whose two y (or two x) edges lie on the same saturated side of erf
// contributes exactly 0 and is skipped: the sum is the same bit for bit, and
// a wave whose voxels all skip a pair skips its four erff evaluations.
// When the pairs are sorted by y (both edges non-decreasing, as MLC banks
// are) and dv > 0, the projected edges are monotone in the pair index, so the
// pairs before the first with y2 > -4.5 and from the first with y1 >= 4.5 on
// are all saturated on one side: two binary searches with the sum's own
// arithmetic bound the window, and the loop inside it is unchanged.
__device__ static inline float head_transmission(const dh_terma_args &a, float hx, float hy,
                                                 float hz, float Z, float inv)
{
    const float dv = (Z - hz) / (a.d_mlc - hz);
    int q0 = 0, q1 = a.n_pairs;
    if (a.sorted && dv > 0.0f) {
        int lo = 0, hi = a.n_pairs;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (pair_y(a.mlc[mid].w, hy, dv, inv) > -kErfSaturated) hi = mid; else lo = mid + 1;
        }
        q0 = lo;
        hi = a.n_pairs;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (pair_y(a.mlc[mid].z, hy, dv, inv) >= kErfSaturated) hi = mid; else lo = mid + 1;
        }
        q1 = lo;
    }
    float tr = 0.0f;
    // Consecutive pairs with equal leaf tips (the same x factor) whose y
    // ranges meet edge to edge form a run, and the run's terms telescope:
    // sum_q ex * (erf(y2_q) - erf(y1_q)) / 2 = ex * (erf(y2_last) - erf(y1_first)) / 2,
    // so a run costs two erf evaluations (one when the previous run ends on
    // its lower edge) instead of one per pair; a rectangular aperture is one
    // run.  Pairs whose edges lie on one saturated side of erf, or whose x
    // factor is exactly 0 (closed), add nothing and end the run.
    bool run = false;
    float run_ex = 0.0f, run_f1 = 0.0f, run_y2 = 0.0f, run_w = NAN, run_x1 = NAN, run_x2 = NAN;
    float prev_w = NAN, prev_erf = 0.0f;            // upper edge of the last run and its erf
    float prev_x1 = NAN, prev_x2 = NAN, prev_ex = 0.0f;
    for (int q = q0; q < q1; q++) {
        const float4 m = a.mlc[q];
        const float e1 = ((m.z - hy) * dv) + hy;
        const float e2 = ((m.w - hy) * dv) + hy;
        const float y1 = e1 * inv, y2 = e2 * inv;
        const bool y_zero = erf_pair_zero(y1, y2);
        if (!y_zero && run && m.x == run_x1 && m.y == run_x2 && m.z == run_w) {
            run_y2 = y2;
            run_w = m.w;
            continue;
        }
        if (run) {
            const float f2 = erff(run_y2);
            tr += run_ex * (0.5f * (f2 - run_f1));
            prev_w = run_w;
            prev_erf = f2;
            run = false;
        }
        if (y_zero) {
            prev_w = NAN;
            continue;
        }
        const float t1 = ((m.x - hx) * dv) + hx;
        const float t2 = ((m.y - hx) * dv) + hx;
        const float x1 = t1 * inv, x2 = t2 * inv;
        if (erf_pair_zero(x1, x2)) {
            prev_w = NAN;
            continue;
        }
        float ex;
        if (m.x == prev_x1 && m.y == prev_x2) {
            ex = prev_ex;
        } else {
            ex = 0.5f * (erff(x2) - erff(x1));
            prev_x1 = m.x;
            prev_x2 = m.y;
            prev_ex = ex;
        }
        if (ex == 0.0f) {
            prev_w = NAN;
            continue;
        }
        run = true;
        run_ex = ex;
        run_x1 = m.x;
        run_x2 = m.y;
        run_w = m.w;
        run_y2 = y2;
        run_f1 = m.z == prev_w ? prev_erf : erff(y1);
    }
    if (run) tr += run_ex * (0.5f * (erff(run_y2) - run_f1));
    return tr;
}

// TERMA and electron contamination (IMRTClasses.cu:209-250) of voxel v =
// (i, j, k) at radiological depth W (cm).
__device__ static inline void terma_voxel(const dh_terma_args &a, uint32_t v, int i, int j, int k, float W,
                                          float *__restrict__ terma, float *__restrict__ electron)
{
    const dh_geom &g = a.g;
    float x, y, z, hx, hy, hz;
    dh_vox_xyz(g, i, j, k, x, y, z);
    dh_img_to_head(g, x, y, z, hx, hy, hz);
    const float dps = a.sad - hz, dss = a.ssd_s - hz;
    // offAxisFactors (IMRTClasses.cu:152-171)
    const float rc = hypotf(hx, hy) * dps / a.sad;
    int lo = 0, hi = a.n_prof;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (rc <= a.prof_r[m]) hi = m; else lo = m + 1;
    }
    float oaf, soft;
    if (lo == a.n_prof) { oaf = a.prof_i[lo - 1]; soft = a.prof_s[lo - 1]; }
    else if (lo == 0) { oaf = a.prof_i[0]; soft = a.prof_s[0]; }
    else {
        const float f = (rc - a.prof_r[lo - 1]) / (a.prof_r[lo] - a.prof_r[lo - 1]);
        oaf = fmaf(a.prof_i[lo] - a.prof_i[lo - 1], f, a.prof_i[lo - 1]);
        soft = fmaf(a.prof_s[lo] - a.prof_s[lo - 1], f, a.prof_s[lo - 1]);
    }
    const float Tp = head_transmission(a, hx, hy, hz, a.sad, a.inv_p);
    const float Ts = head_transmission(a, hx, hy, hz, a.ssd_s, a.inv_s_src);
    const float tr = fminf(1.0f, Ts / Tp);
    const float ip = a.sad / dps, is = a.ssd_s / dss;
    const float fp = (1.0f - tr) * (ip * ip), fs = tr * is;
    float t = 0.0f;
    for (int e = 0; e < a.n_spec; e++) {
        const float att = expf(-a.mu_e[e] * W * soft);
        t += a.pw[e] * att * fp + a.sw[e] * att * fs;
    }
    const float el = fmaxf(0.0f, (expf(-a.el_att * W) - a.el_norm_a) / a.el_norm_b);
    terma[v] = oaf * ((1.0f - a.ssw) * Tp * t + a.ssw * Ts * t);
    electron[v] = a.esw * (0.4f + (0.3f * tr)) * el * Tp;
}

__global__ void __launch_bounds__(kBlock)
k_terma(dh_terma_args a, const float *__restrict__ wet, float *__restrict__ terma,
        float *__restrict__ electron)
{
    const dh_geom &g = a.g;
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nvox) return;
    const int i = (int)(v % (uint32_t)g.nx);
    const uint32_t r = v / (uint32_t)g.nx;
    terma_voxel(a, v, i, (int)(r % (uint32_t)g.ny), (int)(r / (uint32_t)g.ny), wet[v], terma, electron);
}

// {TERMA, density} of the padded grid (border TERMA 0, density air) in one
// of two layouts:
//   pairs (8 B per padded element e):   {T(e), rho(e)}; an x-row of a cell is
//                                       the 16 bytes at e, e + 1
//   quads (32 B per padded element e):  {T, T(x+1) - T} of rows y and y + 1, then
//                                       the same for rho
// The quads' x interpolation fmaf(w, v1 - v0, v0) reads v1 - v0 as stored,
// the same FP32 difference it would form itself, so both layouts give
// bit-identical samples.  Elements of the last padded row / column are never
// read.  (A 16-byte pairs record with stored differences was measured: the
// doubled footprint costs more in cache misses than the saved subtractions.)
__device__ static inline float2 tr_at(const float *__restrict__ terma, const float *__restrict__ dens, int nx,
                                      int ny, int nz, int i, int j, int k)
{
    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) {
        const uint32_t s = (uint32_t)i + (uint32_t)nx * ((uint32_t)j + (uint32_t)ny * (uint32_t)k);
        return make_float2(terma[s], dens[s]);
    }
    return make_float2(0.0f, DH_AIR_DENSITY);
}

template <bool Quad>
__device__ static inline void pack_tr_body(const float *__restrict__ terma, const float *__restrict__ dens, int nx,
                                           int ny, int nz, const dh_pad &p, void *__restrict__ outv)
{
    const uint32_t n = (uint32_t)(p.pxy * p.pz);
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const int i = (int)(t % (uint32_t)p.px) - 1;
        const uint32_t r = t / (uint32_t)p.px;
        const int j = (int)(r % (uint32_t)p.py) - 1;
        const int k = (int)(r / (uint32_t)p.py) - 1;
        const float2 a = tr_at(terma, dens, nx, ny, nz, i, j, k), b = tr_at(terma, dens, nx, ny, nz, i + 1, j, k);
        if (!Quad) {
            ((float2 *)outv)[t] = a;
        } else {
            const float2 c = tr_at(terma, dens, nx, ny, nz, i, j + 1, k);
            const float2 d = tr_at(terma, dens, nx, ny, nz, i + 1, j + 1, k);
            float4 *o = (float4 *)outv + 2 * (size_t)t;
            o[0] = make_float4(a.x, b.x - a.x, c.x, d.x - c.x);
            o[1] = make_float4(a.y, b.y - a.y, c.y, d.y - c.y);
        }
    }
}

template <bool Quad>
__global__ void __launch_bounds__(kBlock)
k_pack_tr(const float *__restrict__ terma, const float *__restrict__ dens, int nx, int ny, int nz,
          dh_pad p, void *__restrict__ outv)
{
    pack_tr_body<Quad>(terma, dens, nx, ny, nz, p, outv);
}

// Bytes per padded element of a layout.
template <bool Quad> constexpr uint32_t tr_elem_bytes() { return Quad ? 32u : 8u; }

// Base addresses of the four x-rows of a trilinear cell, relative to the
// cell's element: rows (y, z), (y + 1, z), (y, z + 1), (y + 1, z + 1).  In the
// quad layout the y + 1 row sits in the same record as row y.
struct tr_bases {
    const char *r00, *r10, *r01, *r11;
};

template <bool Quad>
__device__ static inline tr_bases tr_make_bases(const void *tr, const dh_pad &p)
{
    const size_t eb = tr_elem_bytes<Quad>();
    tr_bases B;
    B.r00 = (const char *)tr;
    B.r10 = B.r00 + (Quad ? 0 : eb * (size_t)p.px);
    B.r01 = B.r00 + eb * (size_t)p.pxy;
    B.r11 = B.r01 + (Quad ? 0 : eb * (size_t)p.px);
    return B;
}

// x-interpolated {TERMA, density} of the cell's four x-rows at weight wx;
// `off` is the byte offset of the cell's element (a 32-bit offset keeps the
// address arithmetic in 32 bits against the rows' uniform bases).
template <bool Quad, typename Off>
__device__ static inline void tr_xrows(const tr_bases &B, Off off, float wx, float &t00, float &t10, float &t01,
                                       float &t11, float &r00, float &r10, float &r01, float &r11)
{
    if (!Quad) {
        // {T, rho} at e and e + 1 of each row (the compiler merges each pair
        // into one 16-byte load)
        const float2 *a = (const float2 *)(B.r00 + off), *b = (const float2 *)(B.r10 + off);
        const float2 *c = (const float2 *)(B.r01 + off), *d = (const float2 *)(B.r11 + off);
        const float2 a0 = a[0], a1 = a[1], b0 = b[0], b1 = b[1], c0 = c[0], c1 = c[1], d0 = d[0], d1 = d[1];
        t00 = fmaf(wx, a1.x - a0.x, a0.x); t10 = fmaf(wx, b1.x - b0.x, b0.x);
        t01 = fmaf(wx, c1.x - c0.x, c0.x); t11 = fmaf(wx, d1.x - d0.x, d0.x);
        r00 = fmaf(wx, a1.y - a0.y, a0.y); r10 = fmaf(wx, b1.y - b0.y, b0.y);
        r01 = fmaf(wx, c1.y - c0.y, c0.y); r11 = fmaf(wx, d1.y - d0.y, d0.y);
    } else {
        const float4 tl = *(const float4 *)(B.r00 + off), rl = *(const float4 *)(B.r00 + off + 16);
        const float4 th = *(const float4 *)(B.r01 + off), rh = *(const float4 *)(B.r01 + off + 16);
        t00 = fmaf(wx, tl.y, tl.x); t10 = fmaf(wx, tl.w, tl.z);
        t01 = fmaf(wx, th.y, th.x); t11 = fmaf(wx, th.w, th.z);
        r00 = fmaf(wx, rl.y, rl.x); r10 = fmaf(wx, rl.w, rl.z);
        r01 = fmaf(wx, rh.y, rh.x); r11 = fmaf(wx, rh.w, rh.z);
    }
}

// Byte offset of padded element (i + 1, j + 1, k + 1) for i, j, k >= -1:
// 32-bit (24-bit products; the host admits volumes under 2 GiB with
// px * py < 2^23) or 64-bit.
template <bool Quad, bool Off32>
__device__ static inline auto tr_offset(const dh_pad &p, int i, int j, int k)
{
    if constexpr (Off32) {
        const int pxy = (int)p.pxy;
        const int e = __mul24(k, pxy) + __mul24(j, p.px) + i + (1 + p.px + pxy);
        return (uint32_t)e * tr_elem_bytes<Quad>();
    } else {
        return (uint64_t)dh_pad_index(p, i, j, k) * tr_elem_bytes<Quad>();
    }
}

__global__ void k_zero_int(int *p, int n)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

// TERMA of the voxel centre sampled like the reference's texture fetch
// (IMRTClasses.cu:267-273).
template <bool Quad>
__device__ static inline float center_terma(const dh_geom &g, const dh_pad &p,
                                            const void *__restrict__ tr, int i, int j, int k)
{
    float x, y, z;
    dh_vox_xyz(g, i, j, k, x, y, z);
    const float tx = fmaf(g.inv_s, x + g.iso[0], 0.5f) - 0.5f;
    const float ty = fmaf(g.inv_s, y + g.iso[1], 0.5f) - 0.5f;
    const float tz = fmaf(g.inv_s, z + g.iso[2], 0.5f) - 0.5f;
    const float ux = __builtin_amdgcn_fmed3f(tx, -1.0f, (float)g.nx);
    const float uy = __builtin_amdgcn_fmed3f(ty, -1.0f, (float)g.ny);
    const float uz = __builtin_amdgcn_fmed3f(tz, -1.0f, (float)g.nz);
    const float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
    const float ax = ux - fx, ay = uy - fy, az = uz - fz;
    float a00, a10, a01, a11, r00, r10, r01, r11;
    tr_xrows<Quad>(tr_make_bases<Quad>(tr, p), tr_offset<Quad, false>(p, (int)fx, (int)fy, (int)fz), ax, a00, a10,
                   a01, a11, r00, r10, r01, r11);
    const float b0 = fmaf(ay, a10 - a00, a00), b1 = fmaf(ay, a11 - a01, a01);
    return fmaf(az, b1 - b0, b0);
}

// Builds the list of voxels with TERMA > 0.01 and zeroes the dose of the rest.
// One atomic per workgroup: waves ballot, the workgroup prefix-sums the four
// wave counts in LDS and reserves its slice of the list at once.
template <bool Quad>
__global__ void __launch_bounds__(kBlock)
k_ccc_compact(dh_geom g, dh_pad p, const void *__restrict__ tr, int *__restrict__ list,
              int *__restrict__ counter, float *__restrict__ dose)
{
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    dh_block_compact<kCompactRounds>(blockIdx.x * (uint32_t)(kBlock * kCompactRounds), nvox, [&](uint32_t v) {
        const int i = (int)(v % (uint32_t)g.nx);
        const uint32_t r = v / (uint32_t)g.nx;
        const int j = (int)(r % (uint32_t)g.ny);
        const int k = (int)(r / (uint32_t)g.ny);
        const float T0 = center_terma<Quad>(g, p, tr, i, j, k);
        const bool act = !(T0 <= 0.01f);
        if (!act) dose[v] = 0.0f;
        return act;
    }, list, counter);
}

// Collapsed-cone superposition (IMRTClasses.cu:252-350) for the listed voxels:
// 6 zenith rows x 12 azimuths; each cone direction is tilted towards the
// source (kernelTilt, IMRTClasses.cu:173-203) and marched from ray_length down
// to 0 in FP32 steps of spacing/10 cm (the reference's own step lattice), with
// the primary (exponential) and scatter (linear) recursions.  Cone results are
// summed with Neumaier compensation.
// Work split: a work item is 64 consecutive listed voxels (one per lane) x
// one of the 12 azimuths, marched through all six zenith rows (equal work).
// All 64 lanes then sample the same cone direction from neighbouring voxels,
// so a wave's gathers touch few cache lines.  The items are short so that the
// last round of waves is nearly full: every wave takes an equal run of items
// and all of them are resident at once, so the kernel lasts ceil(items /
// waves) items.  The loop is co-limited by the texture address/data pipe (16
// cycles per 16-byte wave64 load) and VALU.  Each azimuth's compensated sum
// goes to its own partial buffer and k_ccc_finish adds the 12 in a fixed
// order, so the result is deterministic.
constexpr int kVoxPerItem = 64;
constexpr int kGroups = 12;

// Row items: for voxel lists of at most `cap` voxels (the surrogate's check
// lists, whose launches are short), an item is one zenith row of one azimuth
// (six times more, six times shorter items, so the list's work spreads over
// the chip); each row's result goes to its own partial (stride cap) and
// k_ccc_finish adds an azimuth's six rows in the same compensated order as
// an azimuth item does, so the dose is bit for bit the same.  The choice is
// made on the device from the list's length (uniform over the grid).
// One work item for the voxel v (one lane of the item's 64): azimuth grp,
// zenith rows row0 .. row1 - 1, the rows' results summed with Neumaier
// compensation (rows: a single row, returned as is).  R supplies the rows'
// kernel parameters (nsteps, L0, a, b, c1, c2p, bp, a2) and sp; dirf(cone,
// dx, dy, dz) the cone direction in image space; (ax, ay, az) is the unit
// vector from the isocentre to the source.
template <bool Quad, bool Off32, class R, class Dir>
__device__ static inline float ccc_item(const R &a, const dh_geom &g, const dh_pad &p, const tr_bases &B, float ax,
                                        float ay, float az, uint32_t v, int grp, int row0, int row1, bool rows,
                                        Dir dirf)
{
    const float fnx = (float)g.nx, fny = (float)g.ny, fnz = (float)g.nz;
    const int i = (int)(v % (uint32_t)g.nx);
    const uint32_t r = v / (uint32_t)g.nx;
    const int j = (int)(r % (uint32_t)g.ny);
    const int k = (int)(r / (uint32_t)g.ny);
    float x, y, z, ux, uy, uz;
    dh_vox_xyz(g, i, j, k, x, y, z);
    dh_uvec(g, x, y, z, ux, uy, uz);
    // texel-space origin of the voxel (u = t - 0.5)
    const float ucx = fmaf(g.inv_s, x + g.iso[0], 0.5f) - 0.5f;
    const float ucy = fmaf(g.inv_s, y + g.iso[1], 0.5f) - 0.5f;
    const float ucz = fmaf(g.inv_s, z + g.iso[2], 0.5f) - 0.5f;
    // Rodrigues tilt taking the iso->source axis onto u
    const float c = ax * ux + ay * uy + az * uz;
    const float kx = ay * uz - az * uy;
    const float ky = az * ux - ax * uz;
    const float kz = ax * uy - ay * ux;
    const float inv1c = 1.0f / (1.0f + c);
    float ds = 0.0f, dcmp = 0.0f;
    for (int row = row0; row < row1; row++) {
        const float c1 = a.c1[row], c2p = a.c2p[row], bp = a.bp[row], a2 = a.a2[row];
        const float am = a.a[row], bm = a.b[row], L0 = a.L0[row];
        const int ns = a.nsteps[row];
        {
            float dx0, dy0, dz0;
            dirf(row * 12 + grp, dx0, dy0, dz0);
            const float sc = (kx * dx0 + ky * dy0 + kz * dz0) * inv1c;
            const float dx = (ky * dz0 - kz * dy0) + dx0 * c + kx * sc;
            const float dy = (kz * dx0 - kx * dz0) + dy0 * c + ky * sc;
            const float dz = (kx * dy0 - ky * dx0) + dz0 * c + kz * sc;
            const float tdx = dx * g.inv_s, tdy = dy * g.inv_s, tdz = dz * g.inv_s;
            float Rp = 0.0f, Rs = 0.0f;
            float L = L0;
            for (int st = 0; st < ns; st++) {
                const float Lm = L * 10.0f;
                float qx = __builtin_amdgcn_fmed3f(fmaf(tdx, Lm, ucx), -1.0f, fnx);
                float qy = __builtin_amdgcn_fmed3f(fmaf(tdy, Lm, ucy), -1.0f, fny);
                float qz = __builtin_amdgcn_fmed3f(fmaf(tdz, Lm, ucz), -1.0f, fnz);
                const float fx = floorf(qx), fy = floorf(qy), fz = floorf(qz);
                const float wx = qx - fx, wy = qy - fy, wz = qz - fz;
                float t00, t10, t01, t11, r00, r10, r01, r11;
                tr_xrows<Quad>(B, tr_offset<Quad, Off32>(p, (int)fx, (int)fy, (int)fz), wx, t00, t10, t01, t11, r00,
                               r10, r01, r11);
                const float t0 = fmaf(wy, t10 - t00, t00), t1 = fmaf(wy, t11 - t01, t01);
                const float r0 = fmaf(wy, r10 - r00, r00), r1 = fmaf(wy, r11 - r01, r01);
                const float T = fmaf(wz, t1 - t0, t0);
                const float rho = fmaxf(fmaf(wz, r1 - r0, r0), DH_AIR_DENSITY);
                const float ex = __builtin_amdgcn_exp2f(a2 * rho);       // exp(-a * rho * sp)
                const float K = T * c1;
                Rp = fmaf(Rp - K, ex, K);                                  // Rp e + K (1 - e)
                Rs = fmaf(rho, fmaf(T, c2p, -bp * Rs), Rs);               // Rs (1 - bD) + T D c2
                L = L - a.sp;
            }
            if (rows)
                ds = fmaf(am, Rp, bm * Rs);
            else
                dh_neumaier(ds, dcmp, fmaf(am, Rp, bm * Rs));
        }
    }
    return ds + dcmp;
}

template <bool Quad, bool Off32, bool RowItems = false>
__global__ void __launch_bounds__(kBlock)
k_ccc(dh_ccc_args a, const void *__restrict__ tr, const int *__restrict__ list, const int *__restrict__ counter,
      float *__restrict__ part, uint32_t pstride_az, uint32_t cap)
{
    const dh_geom &g = a.g;
    const tr_bases B = tr_make_bases<Quad>(tr, a.p);
    const int n = *counter;
    const bool Rows = RowItems && (uint32_t)n <= cap;
    const int kItems = Rows ? 6 * kGroups : kGroups;
    const uint32_t pstride = Rows ? cap : pstride_az;
    const int lane = __lane_id();
    // Each wave takes one contiguous run of work items (a chunk's 4 groups are
    // neighbours), and the waves of one XCD take neighbouring runs, so an
    // XCD's gathers stay in one slab of the volume (its own L2).
    const int wave = dh_xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * (kBlock / 64);
    const int nitem = (n + kVoxPerItem - 1) / kVoxPerItem * kItems;
    const int c0 = (int)((int64_t)nitem * wave / nwaves);
    const int c1 = (int)((int64_t)nitem * (wave + 1) / nwaves);
    for (int item = c0; item < c1; item++) {
        const int slot = item % kItems;          // Rows: row * 12 + azimuth
        const int grp = slot % kGroups;
        const int base = item / kItems * kVoxPerItem;
        const int idx = base + lane;
        const bool valid = idx < n;
        const uint32_t v = (uint32_t)list[valid ? idx : base];
        const int row0 = Rows ? slot / kGroups : 0, row1 = Rows ? row0 + 1 : 6;
        const float res = ccc_item<Quad, Off32>(a, g, a.p, B, a.ax, a.ay, a.az, v, grp, row0, row1, Rows,
                                                [&](int cone, float &dx, float &dy, float &dz) {
                                                    dx = a.dir[3 * cone + 0];
                                                    dy = a.dir[3 * cone + 1];
                                                    dz = a.dir[3 * cone + 2];
                                                });
        if (valid) part[(size_t)slot * pstride + idx] = res;
    }
}

// The cone superposition of several control points' voxel lists (one grid,
// pairs layout) in one launch: the points' work items (as in k_ccc, each
// point taking row items when its list has at most cap voxels) are laid end
// to end, so short lists share the chip instead of leaving it idle.  A point's axis and cone directions
// are formed from its geometry with the host's FP32 (dh_uvec, dh_head_to_img
// of the head-frame directions), so each voxel's value equals k_ccc's.
// the struct travels as a kernel argument (at most 4 KiB)
static_assert(sizeof(dh_ccc_multi) <= 4096, "cone arguments of a multi-point launch exceed the kernel argument space");

template <bool Off32>
__global__ void __launch_bounds__(kBlock)
k_ccc_multi(dh_ccc_multi m)
{
    // The dispatcher hands workgroup b to XCD b % 8.  XCD x takes the x-th
    // eighth (slab) of every point's items, the points' slabs end to end,
    // and its waves go through that sequence in rounds of consecutive items:
    // an XCD's gathers stay in one slab of one point's volume at a time (its
    // L2), as a launch per point with one slab per XCD would keep them.  The
    // grid has at least 8 workgroups, so every XCD has waves.
    __shared__ int s_first[DH_CCC_MULTI_MAX + 1], s_n[DH_CCC_MULTI_MAX], s_lo[DH_CCC_MULTI_MAX];
    const int x = blockIdx.x % DH_NUM_XCD;
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int q = 0; q < m.n; q++) {
            const int nq = *m.counter[q];
            const int items = (nq + kVoxPerItem - 1) / kVoxPerItem * ((uint32_t)nq <= m.cap ? 6 * kGroups : kGroups);
            const int lo = (int)((int64_t)items * x / DH_NUM_XCD), hi = (int)((int64_t)items * (x + 1) / DH_NUM_XCD);
            s_first[q] = tot;
            s_n[q] = nq;
            s_lo[q] = lo;
            tot += hi - lo;
        }
        s_first[m.n] = tot;
    }
    __syncthreads();
    const int lane = __lane_id();
    const int nbx = (int)gridDim.x / DH_NUM_XCD + (x < (int)gridDim.x % DH_NUM_XCD ? 1 : 0);
    const int wx = nbx * (kBlock / 64);                                   // waves on this XCD
    const int lw = (int)(blockIdx.x / DH_NUM_XCD) * (kBlock / 64) + (threadIdx.x >> 6);
    const int nseq = s_first[m.n];
    int q = 0;
    for (int i = lw; i < nseq; i += wx) {
        while (i >= s_first[q + 1]) q++;
        // the point is uniform over the wave: keep its arguments in scalars
        q = __builtin_amdgcn_readfirstlane(q);
        const int item = __builtin_amdgcn_readfirstlane(s_lo[q] - s_first[q]) + i;
        const int n = __builtin_amdgcn_readfirstlane(s_n[q]);
        const bool rows = (uint32_t)n <= m.cap;
        const int kItems = rows ? 6 * kGroups : kGroups;
        const uint32_t pstride = rows ? m.cap : m.nvox;
        const int slot = item % kItems;
        const int grp = slot % kGroups;
        const int base = item / kItems * kVoxPerItem;
        const int idx = base + lane;
        const bool valid = idx < n;
        const dh_geom g = m.g[q];
        const uint32_t v = (uint32_t)m.list[q][valid ? idx : base];
        float ax, ay, az;
        dh_uvec(g, 0.0f, 0.0f, 0.0f, ax, ay, az);
        const int row0 = rows ? slot / kGroups : 0, row1 = rows ? row0 + 1 : 6;
        const float res = ccc_item<false, Off32>(m, g, m.p, tr_make_bases<false>(m.tr[q], m.p), ax, ay, az, v, grp,
                                                 row0, row1, rows, [&](int cone, float &dx, float &dy, float &dz) {
                                                     dh_head_to_img(g, m.hdir[3 * cone + 0], m.hdir[3 * cone + 1],
                                                                    m.hdir[3 * cone + 2], dx, dy, dz);
                                                 });
        if (valid) m.part[q][(size_t)slot * pstride + idx] = res;
    }
}

// Dose of the listed voxels from the 12 azimuth sums, the electron term and the
// MU calibration (IMRTClasses.cu:338-346).
__device__ static inline void ccc_finish(float mu_cal, float mu, const float *__restrict__ electron,
                                         const int *__restrict__ list, int n, const float *__restrict__ part,
                                         uint32_t pstride_az, uint32_t cap, float *__restrict__ dose)
{
    const bool Rows = (uint32_t)n <= cap;
    const uint32_t pstride = Rows ? cap : pstride_az;
    for (int idx = blockIdx.x * kBlock + threadIdx.x; idx < n; idx += gridDim.x * kBlock) {
        const uint32_t v = (uint32_t)list[idx];
        float p[kGroups];
        if (Rows) {
#pragma unroll
            for (int q = 0; q < kGroups; q++) {
                float ds = 0.0f, dcmp = 0.0f;
#pragma unroll
                for (int row = 0; row < 6; row++) dh_neumaier(ds, dcmp, part[(size_t)(row * kGroups + q) * pstride + idx]);
                p[q] = ds + dcmp;
            }
        } else {
#pragma unroll
            for (int q = 0; q < kGroups; q++) p[q] = part[(size_t)q * pstride + idx];
        }
        const float s0 = (p[0] + p[1]) + (p[2] + p[3]);
        const float s1 = (p[4] + p[5]) + (p[6] + p[7]);
        const float s2 = (p[8] + p[9]) + (p[10] + p[11]);
        const float dv = ((s0 + s1) + s2) + electron[v];
        dose[v] = (!isnan(dv) && dv >= 0.0f) ? mu_cal * dv * mu : 0.0f;
    }
}

__global__ void __launch_bounds__(kBlock)
k_ccc_finish(float mu_cal, float mu, const float *__restrict__ electron, const int *__restrict__ list,
             const int *__restrict__ counter, const float *__restrict__ part, uint32_t pstride_az, uint32_t cap,
             float *__restrict__ dose)
{
    ccc_finish(mu_cal, mu, electron, list, *counter, part, pstride_az, cap, dose);
}

// k_ccc_finish of each point of a k_ccc_multi launch (blockIdx.y = point).
__global__ void __launch_bounds__(kBlock)
k_ccc_finish_multi(dh_ccc_multi m)
{
    const int q = blockIdx.y;
    ccc_finish(m.mu_cal, m.mu[q], m.electron[q], m.list[q], *m.counter[q], m.part[q], m.nvox, m.cap, m.out[q]);
}

// k_pack_tr (pairs layout) of each point's TERMA with the shared density
// (blockIdx.y = point).
__global__ void __launch_bounds__(kBlock)
k_pack_tr_multi(dh_ccc_multi m, const float *__restrict__ dens)
{
    const int q = blockIdx.y;
    pack_tr_body<false>(m.terma[q], dens, m.g[q].nx, m.g[q].ny, m.g[q].nz, m.p, (void *)m.tr[q]);
}

// ---- radiological depth on a beam's-eye-view grid (checked surrogate path) --
// Fan lines from the source through an a x b lattice on the isocentre plane
// (the plane through the isocentre normal to the central axis); along each
// line the density is sampled every h mm of distance t from the source and
// integrated with the trapezoid rule (zero outside the volume), so that
// table(a, b, k) is the line integral from the source to distance t0 + k h.
// The table is stored in 8 (a) x 4 (k) tiles of 128 bytes, tile rows along a
// for each (b, k / 4): voxels next to each other in the image map to nearby
// (a, k) whatever the beam angle, so a wave's lookups touch few cache lines.
// A voxel's depth is that table interpolated trilinearly at its own (a, b, t):
// one sample per line step instead of one per voxel per step, off the parity
// path (the result is judged by gamma against the plain CCCS).  An entry whose
// sample point lies outside the box of voxel centres is stored as
// -1 - integral; a lattice line takes part in a voxel's lateral interpolation
// only if one of its two samples around the voxel's distance lies inside that
// box, so voxels next to a face parallel to the beam are not averaged with
// lines that run outside the volume, or inside its last half voxel where the
// density blends with the border.
struct bev_grid {
    float e1[3], e2[3], c[3];    // lateral unit vectors and central axis (image frame)
    float sad;                   // |source - isocentre|
    float a0, b0, da;            // lattice origin and pitch on the isocentre plane (mm)
    float t0, h;                 // first distance from the source and step (mm)
    int na, nb, nt;
    int nat, ntt;                // tiles along a (8 lines) and k (4 steps)
};

__device__ static inline size_t bev_index(const bev_grid &b, int ia, int ib, int ik)
{
    return (((size_t)ib * b.ntt + (ik >> 2)) * b.nat + (ia >> 3)) * 32 + (ik & 3) * 8 + (ia & 7);
}

// A workgroup owns 64 consecutive lines (one per lane); its kBevSeg waves
// take consecutive segments of kBevLen steps of those lines.  Each wave keeps
// its segment's running trapezoid sums in registers, the segment totals give
// every segment its offset through LDS, and lines longer than kBevSeg
// segments carry their integral into the next round of segments.
constexpr int kBevSeg = 16, kBevLen = 16;

template <bool Off32>
__global__ void __launch_bounds__(64 * kBevSeg)
k_bev_prefix(dh_geom g, dh_pad p, bev_grid b, const float4 *__restrict__ vol, float *__restrict__ table)
{
    __shared__ float s_tot[kBevSeg][64];
    const int nl = b.na * b.nb;
    const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
    const int l = blockIdx.x * 64 + lane;
    const bool line = l < nl;
    const int lc = line ? l : nl - 1;
    const int ia = lc % b.na, ib = lc / b.na;
    const float4 *vol_hi = vol + p.pxy;
    const float pa = fmaf((float)ia, b.da, b.a0), pb = fmaf((float)ib, b.da, b.b0);
    float dx = fmaf(pa, b.e1[0], pb * b.e2[0]) - g.src[0];
    float dy = fmaf(pa, b.e1[1], pb * b.e2[1]) - g.src[1];
    float dz = fmaf(pa, b.e1[2], pb * b.e2[2]) - g.src[2];
    const float inv = rsqrtf(fmaf(dx, dx, fmaf(dy, dy, dz * dz)));
    dx *= inv; dy *= inv; dz *= inv;
    // texel coordinate of the point at distance t: T0 + t dT
    const float T0x = fmaf(g.inv_s, g.src[0] + g.iso[0], 0.5f), dTx = g.inv_s * dx;
    const float T0y = fmaf(g.inv_s, g.src[1] + g.iso[1], 0.5f), dTy = g.inv_s * dy;
    const float T0z = fmaf(g.inv_s, g.src[2] + g.iso[2], 0.5f), dTz = g.inv_s * dz;
    const float fnx = (float)g.nx, fny = (float)g.ny, fnz = (float)g.nz;
    auto sample = [&](int k, bool &in) -> float {
        const float t = fmaf((float)k, b.h, b.t0);
        const float tx = fmaf(dTx, t, T0x), ty = fmaf(dTy, t, T0y), tz = fmaf(dTz, t, T0z);
        float rho = 0.0f;
        // density over the volume's box (texel 0 .. n, the reference's inside
        // test, border blend included); the flag marks points inside the box
        // of voxel centres (texel 0.5 .. n - 0.5)
        if (k >= 0 && k < b.nt && tx >= 0.0f && tx < fnx && ty >= 0.0f && ty < fny && tz >= 0.0f && tz < fnz)
            rho = fmaxf(dh_trilinear_q<Off32>(vol, vol_hi, p, g.nx, g.ny, g.nz, tx, ty, tz), 0.0f);
        in = k >= 0 && k < b.nt && tx >= 0.5f && tx <= fnx - 0.5f && ty >= 0.5f && ty <= fny - 0.5f &&
             tz >= 0.5f && tz <= fnz - 0.5f;
        return rho;
    };
    const float hh = 0.5f * b.h;
    float carry = 0.0f;
    for (int base = 0; base < b.nt; base += kBevSeg * kBevLen) {
        // step k adds the trapezoid over (k - 1, k)
        const int k0 = base + seg * kBevLen;
        float run[kBevLen];
        uint32_t inside = 0;             // bit q: the sample of step k0 + q lies among the voxel centres
        bool in;
        float prev = sample(k0 - 1, in), acc = 0.0f;
#pragma unroll
        for (int q = 0; q < kBevLen; q++) {
            const float rho = sample(k0 + q, in);
            inside |= (uint32_t)in << q;
            acc = fmaf(hh, prev + rho, acc);
            prev = rho;
            run[q] = acc;
        }
        s_tot[seg][lane] = acc;
        __syncthreads();
        float off = carry, all = carry;
#pragma unroll
        for (int q = 0; q < kBevSeg; q++) {
            const float v = s_tot[q][lane];
            off += q < seg ? v : 0.0f;
            all += v;
        }
        __syncthreads();
        carry = all;
        if (line) {
#pragma unroll
            for (int q = 0; q < kBevLen; q++)
                if (k0 + q < b.nt) {
                    const float v = off + run[q];
                    table[bev_index(b, ia, ib, k0 + q)] = (inside >> q) & 1u ? v : -1.0f - v;
                }
        }
    }
}

// Voxels whose interpolated depth is below w_exact cm (where the electron
// contamination term, which falls to zero at the model's fitted dmax, is
// large and steep in the depth) and whose projection onto the isocentre
// plane lies in the field box fb (head frame, mm; the electron term is
// proportional to the primary transmission, which vanishes a few mm outside
// the open leaf pairs) take the reference's own ray trace instead: 1 mm
// samples from the voxel towards the source until the sample leaves the
// volume.  These rays are short, so the exact depth of the build-up region
// costs a small fraction of a whole-volume trace.
template <bool Off32>
__device__ static inline float bev_depth(const dh_geom &g, const dh_pad &p, const bev_grid &b,
                                         const float *__restrict__ table, const float4 *__restrict__ vol,
                                         float w_exact, const float4 &fbox, float sad_h, float x, float y, float z)
{
    const float wx = x - g.src[0], wy = y - g.src[1], wz = z - g.src[2];
    const float wc = fmaf(wx, b.c[0], fmaf(wy, b.c[1], wz * b.c[2]));
    const float lam = b.sad / wc;     // the host checked wc > 0 for the whole volume
    const float a = lam * fmaf(wx, b.e1[0], fmaf(wy, b.e1[1], wz * b.e1[2]));
    const float bb = lam * fmaf(wx, b.e2[0], fmaf(wy, b.e2[1], wz * b.e2[2]));
    const float t = sqrtf(fmaf(wx, wx, fmaf(wy, wy, wz * wz)));
    // lattice coordinates, clamped so that both corners of each axis exist
    const float fa = fminf(fmaxf((a - b.a0) / b.da, 0.0f), (float)b.na - 1.001f);
    const float fb = fminf(fmaxf((bb - b.b0) / b.da, 0.0f), (float)b.nb - 1.001f);
    const float fk = fminf(fmaxf((t - b.t0) / b.h, 0.0f), (float)b.nt - 1.001f);
    const int ia = (int)fa, ib = (int)fb, ik = (int)fk;
    const float ua = fa - (float)ia, ub = fb - (float)ib, uk = fk - (float)ik;
    // per lattice line (a, b corner): the integral at the voxel's distance
    // (linear in k), and whether the line is inside the volume there
    float sum_in = 0.0f, w_in = 0.0f, sum_all = 0.0f;
#pragma unroll
    for (int db = 0; db < 2; db++)
#pragma unroll
        for (int da = 0; da < 2; da++) {
            const float r0 = table[bev_index(b, ia + da, ib + db, ik)];
            const float r1 = table[bev_index(b, ia + da, ib + db, ik + 1)];
            const float v0 = r0 >= 0.0f ? r0 : -1.0f - r0, v1 = r1 >= 0.0f ? r1 : -1.0f - r1;
            const float val = fmaf(uk, v1 - v0, v0);
            const float w = (da ? ua : 1.0f - ua) * (db ? ub : 1.0f - ub);
            sum_all = fmaf(w, val, sum_all);
            if (r0 >= 0.0f || r1 >= 0.0f) {
                sum_in = fmaf(w, val, sum_in);
                w_in += w;
            }
        }
    // lines that miss the volume at this distance do not count, unless
    // (nearly) all four do
    const float d = w_in > 1e-3f ? sum_in / w_in : sum_all;
    float w = 0.1f * d;                  // mm of unit density -> cm
    bool exact = w < w_exact;
    if (exact) {
        float hx, hy, hz;
        dh_img_to_head(g, x, y, z, hx, hy, hz);
        const float mag = sad_h / (sad_h - hz);
        const float xi = hx * mag, yi = hy * mag;
        exact = xi >= fbox.x && xi <= fbox.y && yi >= fbox.z && yi <= fbox.w;
    }
    if (exact) {
        const float4 *vol_hi = vol + p.pxy;
        float ux, uy, uz;
        dh_uvec(g, x, y, z, ux, uy, uz);
        const float fnx = (float)g.nx, fny = (float)g.ny, fnz = (float)g.nz;
        double acc = 0.0;
        float part = 0.0f;
        for (int m = 0;; m++) {
            const float L = (float)m;
            const float tx = fmaf(g.inv_s, fmaf(ux, L, x) + g.iso[0], 0.5f);
            const float ty = fmaf(g.inv_s, fmaf(uy, L, y) + g.iso[1], 0.5f);
            const float tz = fmaf(g.inv_s, fmaf(uz, L, z) + g.iso[2], 0.5f);
            if (!(tx >= 0.0f && tx < fnx && ty >= 0.0f && ty < fny && tz >= 0.0f && tz < fnz)) break;
            part += fmaxf(dh_trilinear_q<Off32>(vol, vol_hi, p, g.nx, g.ny, g.nz, tx, ty, tz), 0.0f);
            if ((m & 3) == 3) {
                acc += (double)part;
                part = 0.0f;
            }
        }
        acc += (double)part;
        w = (float)(-0.05 + 0.1 * acc);
    }
    return w;
}

// The depth of a voxel from the table (bev_depth) goes straight into its
// TERMA (terma_voxel), written to wet_out too when that is not null.
template <bool Off32>
__global__ void __launch_bounds__(kBlock)
k_terma_bev(dh_terma_args a, dh_pad p, bev_grid b, const float *__restrict__ table, const float4 *__restrict__ vol,
            float w_exact, float4 fbox, float sad_h, float *__restrict__ wet_out, float *__restrict__ terma,
            float *__restrict__ electron)
{
    const dh_geom &g = a.g;
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const uint32_t v = blockIdx.x * kBlock + threadIdx.x;
    if (v >= nvox) return;
    const int i = (int)(v % (uint32_t)g.nx);
    const uint32_t r = v / (uint32_t)g.nx;
    const int j = (int)(r % (uint32_t)g.ny), k = (int)(r / (uint32_t)g.ny);
    float x, y, z;
    dh_vox_xyz(g, i, j, k, x, y, z);
    const float w = bev_depth<Off32>(g, p, b, table, vol, w_exact, fbox, sad_h, x, y, z);
    if (wet_out) wet_out[v] = w;
    terma_voxel(a, v, i, j, k, w, terma, electron);
}

}  // namespace

void dh_launch_terma(const dh_terma_args &a, const float *wet, float *terma, float *electron,
                     hipStream_t s)
{
    size_t n = (size_t)a.g.nx * a.g.ny * a.g.nz;
    dh_timed t("terma", s);
    hipLaunchKernelGGL(k_terma, dim3(grid_for(n)), dim3(kBlock), 0, s, a, wet, terma, electron);
    DH_HIP(hipGetLastError());
}

namespace {
// The beam's-eye-view grid of a geometry and its table in table_slot (the
// prefix launch unless table_ready); false (nothing launched) where the grid
// does not apply.
struct bev_setup {
    bev_grid b;
    float *table;
    dh_pad p;
    float4 fbox;
    bool off32;
    size_t nvox;
};
bool bev_table(int dev, const dh_geom &g, const float4 *vol_pad, const float *field_box, int table_slot,
               bool table_ready, hipStream_t s, bev_setup &out)
{
    const float4 fbox = field_box ? make_float4(field_box[0], field_box[1], field_box[2], field_box[3])
                                : make_float4(1.0f, -1.0f, 1.0f, -1.0f);
    const double S[3] = {g.src[0], g.src[1], g.src[2]};
    const double sad = std::sqrt(S[0] * S[0] + S[1] * S[1] + S[2] * S[2]);
    if (!(sad > 1.0)) return false;
    const double c[3] = {-S[0] / sad, -S[1] / sad, -S[2] / sad};
    // lateral basis: c x (the axis least aligned with c), then c x e1
    int ax = 0;
    for (int q = 1; q < 3; q++)
        if (std::fabs(c[q]) < std::fabs(c[ax])) ax = q;
    double ref[3] = {0.0, 0.0, 0.0};
    ref[ax] = 1.0;
    double e1[3] = {c[1] * ref[2] - c[2] * ref[1], c[2] * ref[0] - c[0] * ref[2], c[0] * ref[1] - c[1] * ref[0]};
    const double n1 = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
    for (double &e : e1) e /= n1;
    const double e2[3] = {c[1] * e1[2] - c[2] * e1[1], c[2] * e1[0] - c[0] * e1[2], c[0] * e1[1] - c[1] * e1[0]};
    // the volume's box (voxel centres +- half a voxel, relative to the isocentre)
    const int n[3] = {g.nx, g.ny, g.nz};
    double lo[3], hi[3];
    for (int q = 0; q < 3; q++) {
        lo[q] = -0.5 * g.s - g.iso[q];
        hi[q] = ((double)n[q] - 0.5) * g.s - g.iso[q];
    }
    double amin = 1e300, amax = -1e300, bmin = 1e300, bmax = -1e300, t1 = 0.0;
    for (int corner = 0; corner < 8; corner++) {
        double w[3];
        for (int q = 0; q < 3; q++) w[q] = ((corner >> q) & 1 ? hi[q] : lo[q]) - S[q];
        const double wc = w[0] * c[0] + w[1] * c[1] + w[2] * c[2];
        // every point of the volume must lie well in front of the source
        if (!(wc > 0.05 * sad)) return false;
        const double lam = sad / wc;
        const double a = lam * (w[0] * e1[0] + w[1] * e1[1] + w[2] * e1[2]);
        const double b = lam * (w[0] * e2[0] + w[1] * e2[1] + w[2] * e2[2]);
        amin = std::min(amin, a); amax = std::max(amax, a);
        bmin = std::min(bmin, b); bmax = std::max(bmax, b);
        t1 = std::max(t1, std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));
    }
    double d2 = 0.0;                         // distance from the source to the box
    for (int q = 0; q < 3; q++) {
        const double cq = std::min(std::max(S[q], lo[q]), hi[q]);
        d2 += (S[q] - cq) * (S[q] - cq);
    }
    bev_grid b;
    b.h = g.s;
    b.da = (float)(g.s * sad / t1);          // line pitch <= one voxel everywhere in the volume
    b.a0 = (float)(amin - 2.0 * b.da);
    b.b0 = (float)(bmin - 2.0 * b.da);
    b.na = (int)std::ceil((amax - amin) / b.da) + 5;
    b.nb = (int)std::ceil((bmax - bmin) / b.da) + 5;
    b.t0 = (float)std::max(0.0, std::sqrt(d2) - 2.0 * b.h);
    b.nt = (int)std::ceil((t1 - b.t0) / b.h) + 3;
    b.nat = (b.na + 7) / 8;
    b.ntt = (b.nt + 3) / 4;
    const double cells = (double)b.nat * 8 * b.nb * b.ntt * 4;
    if (!(cells < (double)(1u << 28)) || (double)b.na * b.nb > (double)(1u << 30)) return false;
    for (int q = 0; q < 3; q++) {
        b.e1[q] = (float)e1[q];
        b.e2[q] = (float)e2[q];
        b.c[q] = (float)c[q];
    }
    b.sad = (float)sad;
    float *table = (float *)dh_scratch(dev, table_slot, sizeof(float) * (size_t)cells);
    const dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
    const size_t nl = (size_t)b.na * b.nb, nvox = (size_t)g.nx * g.ny * g.nz;
    const bool off32 = (size_t)16 * p.pxy * p.pz < ((size_t)1 << 31) && p.pxy < ((int64_t)1 << 23);
    if (!table_ready) {
        dh_timed t("bev_prefix", s);
        const unsigned lgrid = (unsigned)((nl + 63) / 64);
        if (off32)
            hipLaunchKernelGGL(k_bev_prefix<true>, dim3(lgrid), dim3(64 * kBevSeg), 0, s, g, p, b, vol_pad, table);
        else
            hipLaunchKernelGGL(k_bev_prefix<false>, dim3(lgrid), dim3(64 * kBevSeg), 0, s, g, p, b, vol_pad, table);
        DH_HIP(hipGetLastError());
    }
    out.b = b;
    out.table = table;
    out.p = p;
    out.fbox = fbox;
    out.off32 = off32;
    out.nvox = nvox;
    return true;
}
}  // namespace

bool dh_launch_bev_terma(int dev, const dh_terma_args &a, const float4 *vol_pad, float w_exact_cm,
                         const float *field_box, float *terma, float *electron, float *wet_out, hipStream_t s,
                         int table_slot, bool table_ready)
{
    bev_setup u;
    if (!bev_table(dev, a.g, vol_pad, field_box, table_slot, table_ready, s, u)) return false;
    dh_timed t("terma", s);
    if (u.off32)
        hipLaunchKernelGGL(k_terma_bev<true>, dim3(grid_for(u.nvox)), dim3(kBlock), 0, s, a, u.p, u.b, u.table,
                           vol_pad, w_exact_cm, u.fbox, a.sad, wet_out, terma, electron);
    else
        hipLaunchKernelGGL(k_terma_bev<false>, dim3(grid_for(u.nvox)), dim3(kBlock), 0, s, a, u.p, u.b, u.table,
                           vol_pad, w_exact_cm, u.fbox, a.sad, wet_out, terma, electron);
    DH_HIP(hipGetLastError());
    return true;
}

int dh_ccc_layout(const dh_ccc_args &a, int choice)
{
    if (choice < 0) choice = dh_tune_lookup("ccc", dh_tune_key_geom(a.g), 0);
    DH_REQUIRE(choice == 0 || choice == 1, "CCCS layout choice must be 0 (pairs) or 1 (quads)");
    return choice;
}

// The packed volume (up to 32 B per padded element), then the cone kernel's
// 12 azimuth sums per voxel.
static size_t tr_vol_bytes(const dh_ccc_args &a) { return ((size_t)32 * a.p.pxy * a.p.pz + 255) & ~(size_t)255; }
size_t dh_tr_pad_bytes(const dh_ccc_args &a)
{
    return tr_vol_bytes(a) + sizeof(float) * kGroups * (size_t)a.g.nx * a.g.ny * a.g.nz;
}

namespace {
// 32-bit byte offsets serve packed volumes under 2 GiB whose padded planes fit
// the 24-bit products (px * py < 2^23); larger ones take 64-bit offsets.
bool ccc_off32(const dh_ccc_args &a)
{
    return tr_vol_bytes(a) < ((size_t)1 << 31) && a.p.pxy < ((int64_t)1 << 23);
}

// Cone kernel + finish over the first *counter entries of list (at most
// max_n, which sizes the grids).
void launch_k_ccc(const char *timer, const dh_ccc_args &a, int layout, int64_t max_n, const void *tr,
                  const float *electron, const int *list, const int *counter, float *dose, hipStream_t s,
                  bool row_items)
{
    const bool o32 = ccc_off32(a);
    const uint32_t nvox = (uint32_t)((size_t)a.g.nx * a.g.ny * a.g.nz);
    DH_REQUIRE(max_n >= 0 && max_n <= (int64_t)nvox, "cone list longer than the volume");
    float *part = (float *)((char *)tr + tr_vol_bytes(a));
    // lists of at most 16384 voxels take row items: their azimuth items
    // (12 per 64 voxels, at most 3072) would leave most of the chip's waves
    // idle, while longer lists fill it either way and the row items' six
    // times more partials only cost traffic there (measured on the checked
    // surrogate's lattice and fallback lists); 72 partials per voxel must fit
    // the 12-per-voxel partial buffer (at most nvox / 6 voxels)
    const uint32_t cap = row_items ? (uint32_t)std::min<int64_t>(std::min<int64_t>(max_n, nvox / 6), 16384) : 0u;
    const int64_t items = (max_n + kVoxPerItem - 1) / kVoxPerItem * kGroups +
                          (cap ? ((int64_t)cap + kVoxPerItem - 1) / kVoxPerItem * kGroups * 5 : 0);
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((items + 3) / 4, 2048));
    {
        dh_timed t(timer, s);
        if (cap && layout && o32)
            hipLaunchKernelGGL((k_ccc<true, true, true>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part,
                               nvox, cap);
        else if (cap && layout)
            hipLaunchKernelGGL((k_ccc<true, false, true>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part,
                               nvox, cap);
        else if (cap && o32)
            hipLaunchKernelGGL((k_ccc<false, true, true>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part,
                               nvox, cap);
        else if (cap)
            hipLaunchKernelGGL((k_ccc<false, false, true>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part,
                               nvox, cap);
        else if (layout && o32)
            hipLaunchKernelGGL((k_ccc<true, true>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part, nvox,
                               cap);
        else if (layout)
            hipLaunchKernelGGL((k_ccc<true, false>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part, nvox,
                               cap);
        else if (o32)
            hipLaunchKernelGGL((k_ccc<false, true>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part, nvox,
                               cap);
        else
            hipLaunchKernelGGL((k_ccc<false, false>), dim3(grid), dim3(kBlock), 0, s, a, tr, list, counter, part, nvox,
                               cap);
        DH_HIP(hipGetLastError());
    }
    const unsigned fgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_n + kBlock - 1) / kBlock, 4096));
    dh_timed t("ccc_finish", s);
    hipLaunchKernelGGL(k_ccc_finish, dim3(fgrid), dim3(kBlock), 0, s, a.mu_cal, a.mu, electron, list, counter,
                       (const float *)part, nvox, cap, dose);
    DH_HIP(hipGetLastError());
}
}  // namespace

void dh_launch_pack_tr(const dh_ccc_args &a, const float *terma, const float *density, void *tr_pad, int layout,
                       hipStream_t s)
{
    const dh_geom &g = a.g;
    size_t np = (size_t)a.p.pxy * a.p.pz;
    const unsigned grid = (unsigned)std::min<size_t>(grid_for(np), 4096);
    dh_timed t("pack_tr", s);
    if (layout)
        hipLaunchKernelGGL(k_pack_tr<true>, dim3(grid), dim3(kBlock), 0, s, terma, density, g.nx, g.ny, g.nz, a.p,
                           tr_pad);
    else
        hipLaunchKernelGGL(k_pack_tr<false>, dim3(grid), dim3(kBlock), 0, s, terma, density, g.nx, g.ny, g.nz, a.p,
                           tr_pad);
    DH_HIP(hipGetLastError());
}

void dh_launch_ccc_counted(const dh_ccc_args &a, const void *tr_pad, int layout, const float *electron,
                           const int *list, const int *counter, int64_t max_n, float *dose, hipStream_t s)
{
    if (max_n <= 0) return;
    launch_k_ccc("ccc_list", a, layout, max_n, tr_pad, electron, list, counter, dose, s, true);
}

void dh_ccc_multi_init(dh_ccc_multi &m, const dh_ccc_args *cas, int n)
{
    DH_REQUIRE(n >= 1 && n <= DH_CCC_MULTI_MAX, "cone launch of 1..8 control points");
    std::memset(&m, 0, sizeof(m));
    const dh_ccc_args &a = cas[0];
    m.p = a.p;
    m.sp = a.sp;
    m.mu_cal = a.mu_cal;
    for (int r = 0; r < 6; r++) {
        m.nsteps[r] = a.nsteps[r];
        m.L0[r] = a.L0[r]; m.a[r] = a.a[r]; m.b[r] = a.b[r];
        m.c1[r] = a.c1[r]; m.c2p[r] = a.c2p[r]; m.bp[r] = a.bp[r]; m.a2[r] = a.a2[r];
    }
    std::memcpy(m.hdir, a.hdir, sizeof(m.hdir));
    m.nvox = (uint32_t)((size_t)a.g.nx * a.g.ny * a.g.nz);
    m.n = n;
    for (int q = 0; q < n; q++) {
        DH_REQUIRE(cas[q].g.nx == a.g.nx && cas[q].g.ny == a.g.ny && cas[q].g.nz == a.g.nz && cas[q].sp == a.sp &&
                   cas[q].mu_cal == a.mu_cal, "the points of a cone launch must share grid and model");
        m.g[q] = cas[q].g;
        m.mu[q] = cas[q].mu;
    }
}

size_t dh_ccc_multi_tr_bytes(const dh_ccc_args &a) { return ((size_t)8 * a.p.pxy * a.p.pz + 255) & ~(size_t)255; }
size_t dh_ccc_multi_part_bytes(const dh_ccc_args &a)
{
    return sizeof(float) * kGroups * (size_t)a.g.nx * a.g.ny * a.g.nz;
}

void dh_launch_pack_tr_multi(const dh_ccc_multi &m, const float *density, hipStream_t s)
{
    const size_t np = (size_t)m.p.pxy * m.p.pz;
    const unsigned grid = (unsigned)std::min<size_t>(grid_for(np), 4096);
    dh_timed t("pack_tr", s);
    hipLaunchKernelGGL(k_pack_tr_multi, dim3(grid, m.n), dim3(kBlock), 0, s, m, density);
    DH_HIP(hipGetLastError());
}

void dh_launch_ccc_multi(const char *timer, dh_ccc_multi m, int64_t max_n, hipStream_t s)
{
    if (max_n <= 0) return;
    DH_REQUIRE(max_n <= (int64_t)m.nvox, "cone list longer than the volume");
    // 32-bit offsets as for k_ccc (the pairs volume is a quarter of the size
    // that test admits)
    const bool o32 = (size_t)32 * m.p.pxy * m.p.pz < ((size_t)1 << 31) && m.p.pxy < ((int64_t)1 << 23);
    // row items for lists of at most 16384 voxels, as launch_k_ccc
    m.cap = (uint32_t)std::min<int64_t>(std::min<int64_t>(max_n, m.nvox / 6), 16384);
    const int64_t items = m.n * ((max_n + kVoxPerItem - 1) / kVoxPerItem * kGroups +
                                 ((int64_t)m.cap + kVoxPerItem - 1) / kVoxPerItem * kGroups * 5);
    const unsigned grid = (unsigned)std::max<int64_t>(DH_NUM_XCD, std::min<int64_t>((items + 3) / 4, 2048));
    {
        dh_timed t(timer, s);
        if (o32)
            hipLaunchKernelGGL(k_ccc_multi<true>, dim3(grid), dim3(kBlock), 0, s, m);
        else
            hipLaunchKernelGGL(k_ccc_multi<false>, dim3(grid), dim3(kBlock), 0, s, m);
        DH_HIP(hipGetLastError());
    }
    const unsigned fgrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((max_n + kBlock - 1) / kBlock, 4096));
    dh_timed t("ccc_finish", s);
    hipLaunchKernelGGL(k_ccc_finish_multi, dim3(fgrid, m.n), dim3(kBlock), 0, s, m);
    DH_HIP(hipGetLastError());
}

void dh_launch_ccc(const dh_ccc_args &a, const float *terma, const float *density,
                   const float *electron, void *tr_pad, int *active, int *counter,
                   float *dose, hipStream_t s, int choice)
{
    const dh_geom &g = a.g;
    const int layout = dh_ccc_layout(a, choice);
    size_t n = (size_t)g.nx * g.ny * g.nz;
    dh_launch_pack_tr(a, terma, density, tr_pad, layout, s);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, counter, 1);
    DH_HIP(hipGetLastError());
    {
        dh_timed t("ccc_compact", s);
        if (layout)
            hipLaunchKernelGGL(k_ccc_compact<true>, dim3(grid_compact(n)), dim3(kBlock), 0, s, g, a.p,
                               (const void *)tr_pad, active, counter, dose);
        else
            hipLaunchKernelGGL(k_ccc_compact<false>, dim3(grid_compact(n)), dim3(kBlock), 0, s, g, a.p,
                               (const void *)tr_pad, active, counter, dose);
        DH_HIP(hipGetLastError());
    }
    launch_k_ccc("ccc", a, layout, (int64_t)n, tr_pad, electron, active, counter, dose, s, false);
}

void dh_launch_ccc_list(const dh_ccc_args &a, const float *terma, const float *density,
                        const float *electron, void *tr_pad, const int *list, int n,
                        int *counter, float *dose, hipStream_t s)
{
    const int layout = dh_ccc_layout(a, -1);
    dh_launch_pack_tr(a, terma, density, tr_pad, layout, s);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, s, counter, n);
    DH_HIP(hipGetLastError());
    if (n <= 0) return;
    dh_launch_ccc_counted(a, tr_pad, layout, electron, list, counter, n, dose, s);
}
