// influence.hip -- batched proton dose-influence matrix in CSR form.
//
// Rows are spots (mu = 1), columns voxel indices (u32), values the pencil-beam
// dose of that spot at that voxel, the same formula as the dense pencil
// kernel (IMPTClasses.cu:204-271), kept where value >= rel * (row peak).
//
// Work decomposition: the spots of a layer are grouped into blocks of S
// spatially adjacent spots; a block's candidate voxels are those whose
// beam's-eye-view position lies in the block's lateral box (spot positions
// +- the truncation radius), swept in image order k, j, i.  The (k, j) rows of
// a block are cut into C chunks; a work item is (block, chunk) and is pulled by
// a persistent kernel from an atomic queue (largest first).  For each voxel of
// a sweep the per-(voxel, layer) factors are computed once and shared by the
// S spots.  Three passes use the same sweep:
//   PEAK  : per-spot maximum (atomicMax on the float bits, values >= 0);
//   COUNT : per (spot, chunk) number of entries >= threshold;
//   FILL  : entries written at offsets from the exclusive scan of the counts,
//           ballot-compacted, so each row's columns come out in ascending
//           voxel order and the output is deterministic.
#include "dh_runtime.h"

#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

enum { INF_PEAK = 0, INF_COUNT = 1, INF_FILL = 2 };

struct dh_infl_block {
    int layer, first, n, pad_;   // first: index of the block's spot list
    float bx0, bx1, by0, by1;    // BEV (head) lateral box
    int k0, k1, j0, j1;          // image rows swept (inclusive)
    float ex0, ex1, ey0, ey1;    // extent of the spot axes (the box without the radius)
    float R, pad1_, pad2_, pad3_;  // largest spot radius of the block (box = extent +- R)
};

struct dh_infl_args {
    dh_geom g;
    float vsadx, vsady;
    int n_items, C;
    const dh_infl_block *blocks;
    const int *block_spots;      // S entries per block (-1 padded)
    const dh_layer_k *layers;
    const float4 *luts;
    const float *spots;          // 8 floats per spot: x, y, t^x, t^y, mu
    const float *thr;            // per spot threshold (COUNT / FILL)
    const float *wet;
    unsigned *peak_bits;         // per spot (PEAK)
    int *counts;                 // per spot * C (COUNT)
    int *spans;                  // per work item: first and last packed chunk with entries
                                 // (COUNT writes, FILL reads; -1: none)
    const int64_t *offsets;      // per spot * C (FILL)
    const int2 *active;          // [layer][nz][ny] first/last i with w <= cut (lo > hi: none)
    const float4 *factors;       // [layer][voxel] {ap, log2 P, ah, log2 H} (P, H >= 0 folded)
    const float4 *row_bounds;    // [layer][nz * ny] maxima of {sigma_t, sigma_ht, P, H} per image row
    uint32_t *col;
    float *val;
    uint32_t nnz32;              // FILL with 32-bit positions: end of the field's slots
    int *queue;
    unsigned long long *stats;   // optional: lane visits, active visits, chunks, rows
};

namespace {

// Interval of i (0 <= i < n) for which lo <= a + b*i <= hi, widened by one
// voxel on each side (the box is only a pre-filter; the threshold decides).
__device__ static inline void interval(float a, float b, float lo, float hi, int n, int &ilo, int &ihi)
{
    if (fabsf(b) < 1e-12f) {
        if (a < lo || a > hi) { ilo = 1; ihi = 0; }
        return;
    }
    float t0 = (lo - a) / b, t1 = (hi - a) / b;
    if (t0 > t1) { float t = t0; t0 = t1; t1 = t; }
    const int l = (int)floorf(fmaxf(t0, -2.0f)) - 1;
    const int h = (int)ceilf(fminf(t1, (float)n + 2.0f)) + 1;
    ilo = max(ilo, max(l, 0));
    ihi = min(ihi, min(h, n - 1));
}

// P32 (FILL): write positions in 32 bits, bounded by the matrix's entry count
// instead of each slot's end (COUNT and FILL evaluate the same FP32
// expressions, so an entry never leaves its slot; the bound keeps every write
// inside the arrays), which keeps the fill's per-spot state in fewer VGPRs.
template <int S, int MODE, bool P32 = false>
__global__ void __launch_bounds__(64)
k_influence(dh_infl_args a)
{
    using PosT = typename std::conditional<P32, uint32_t, int64_t>::type;
    __shared__ int s_item;
    __shared__ int s_rlo[64], s_rex[64], s_rid[64];   // packed rows of a batch
    const dh_geom &g = a.g;
    const int lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    for (;;) {
        if (lane == 0) s_item = atomicAdd(a.queue, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= a.n_items) break;
        const int bi = item / a.C, chunk = item % a.C;
        const dh_infl_block B = a.blocks[bi];
        const dh_layer_k L = a.layers[B.layer];
        // spot records of the block (wave-uniform)
        float sx[S], sy[S], tx[S], ty[S], th[S];
        int sid[S];
        // smallest radius threshold of the block: its entry threshold (COUNT,
        // FILL) or the lower bound of its peak (PEAK)
        float tmin = INFINITY;
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int id = a.block_spots[B.first * S + s];
            sid[s] = id;
            const int c = id >= 0 ? id : 0;
            sx[s] = a.spots[8 * c]; sy[s] = a.spots[8 * c + 1];
            tx[s] = a.spots[8 * c + 2]; ty[s] = a.spots[8 * c + 3];
            const float tq = id < 0 ? INFINITY : a.thr[c];
            th[s] = MODE == INF_PEAK ? INFINITY : tq;
            tmin = fminf(tmin, tq);
        }
        float pk[S];
        int cnt[S];
        PosT cur64[S];
        float lth[S];
#pragma unroll
        for (int s = 0; s < S; s++) {
            pk[s] = 0.0f; cnt[s] = 0;
            cur64[s] = (MODE == INF_FILL && sid[s] >= 0) ? (PosT)a.offsets[(size_t)sid[s] * a.C + chunk] : 0;
            lth[s] = log2f(th[s]) - 1.01f;
        }
        // Packed chunks of an item come in the same order in COUNT and FILL
        // (same blocks, rows and packing), so COUNT records the first and last
        // one where any spot kept entries, and FILL skips the chunks before the
        // first and ends the item after the last.
        int span_first = 0x7fffffff, span_last = -1;
        if (MODE == INF_FILL) {
            const int2 sp = a.spans ? ((const int2 *)a.spans)[item] : make_int2(0, 0x7fffffff);
            span_first = __builtin_amdgcn_readfirstlane(sp.x);
            span_last = __builtin_amdgcn_readfirstlane(sp.y);
        }
        // slot end of each (spot, chunk): no write can leave its slot even if a
        // fill decision disagreed with the count pass
        int64_t lim[S];
#pragma unroll
        for (int s = 0; s < S; s++)
            lim[s] = (MODE == INF_FILL && !P32 && sid[s] >= 0) ? a.offsets[(size_t)sid[s] * a.C + chunk + 1] : 0;
        // rows of this chunk
        const int nj = B.j1 - B.j0 + 1, nk = B.k1 - B.k0 + 1;
        const int64_t nrows = (int64_t)nj * nk;
        const int64_t r0 = nrows * chunk / a.C, r1 = nrows * (chunk + 1) / a.C;
        // head coordinates are linear in i: h(i) = h(0) + i * dh
        float dhx, dhy, dhz;
        {
            float x0, y0, z0, x1, y1, z1;
            dh_img_to_head(g, 0.0f, 0.0f, 0.0f, x0, y0, z0);
            dh_img_to_head(g, g.s, 0.0f, 0.0f, x1, y1, z1);
            dhx = x1 - x0; dhy = y1 - y0; dhz = z1 - z0;
        }
        const int2 *act_rows = a.active + (size_t)B.layer * g.nz * g.ny;
        unsigned long long st_visit = 0, st_act = 0, st_chunk = 0, st_rows = 0;
        int pc = 0;                      // packed chunk index within the item
        bool done = MODE == INF_FILL && span_first > span_last;
        for (int64_t rb = r0; rb < r1 && !done; rb += 64) {
            // lane r prepares row rb + r: box interval intersected with the
            // layer's active interval on that row
            int my_lo = 1, my_hi = 0;
            const int64_t rr = rb + lane;
            if (rr < r1) {
                const int k = B.k0 + (int)(rr / nj), j = B.j0 + (int)(rr % nj);
                const int2 ar = act_rows[(size_t)k * g.ny + j];
                if (ar.x <= ar.y) {
                    float x, y, z, hx0, hy0, hz0;
                    dh_vox_xyz(g, 0, j, k, x, y, z);
                    dh_img_to_head(g, x, y, z, hx0, hy0, hz0);
                    // this row's radius: beyond it both Gaussian terms of every
                    // reached voxel of the row are below tmin / 2 (the row's
                    // maxima of sigma_t, sigma_ht, P and H bound each voxel's)
                    const float4 rb = a.row_bounds[(size_t)B.layer * g.ny * g.nz + (size_t)k * g.ny + j];
                    float R2 = 0.0f;
                    if (2.0f * rb.z > tmin) R2 = 2.0f * rb.x * rb.x * logf(2.0f * rb.z / tmin);
                    if (2.0f * rb.w > tmin) R2 = fmaxf(R2, 2.0f * rb.y * rb.y * logf(2.0f * rb.w / tmin));
                    const float R = fminf(sqrtf(R2) * 1.002f + 0.5f * g.s, B.R);
                    int ilo = ar.x, ihi = ar.y;
                    interval(hx0, dhx, B.ex0 - R, B.ex1 + R, g.nx, ilo, ihi);
                    interval(hy0, dhy, B.ey0 - R, B.ey1 + R, g.nx, ilo, ihi);
                    my_lo = ilo; my_hi = ihi;
                }
            }
            // Pack the batch's non-empty row intervals end to end so that every
            // 64-voxel chunk fills all lanes: compacted row q starts at flat
            // index s_rex[q]; the order (rows ascending, i ascending) keeps each
            // spot's columns ascending.
            const bool ne = my_lo <= my_hi;
            const uint64_t nem = __builtin_amdgcn_ballot_w64(ne);
            if (!nem) continue;
            const int len = ne ? my_hi - my_lo + 1 : 0;
            int incl = len;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const int t = __shfl_up(incl, off, 64);
                if (lane >= off) incl += t;
            }
            const int total = __shfl(incl, 63, 64);
            const int nrow = __popcll(nem);
            if (ne) {
                const int q = __popcll(nem & below);
                s_rlo[q] = my_lo;
                s_rex[q] = incl - len;
                s_rid[q] = lane;
            }
            __syncthreads();
            st_rows += nrow;
            int cur = 0;                 // compacted row of the next chunk's first voxel
            // Voxel of this lane in the chunk at flat position c0: rows starting
            // inside the chunk come from a uniform walk (usually 1-2 rows).
            // The WET and factor loads of chunk n + 1 are issued before chunk n's
            // spots are evaluated, so their latency overlaps that work.
            auto locate = [&](int c0, int &ci, int &cj, int &ck) {
                const int f = min(c0 + lane, total - 1);
                int q = cur, my_q = cur;
                for (int nq = cur + 1; nq < nrow; nq++) {
                    const int e = s_rex[nq];
                    if (e > c0 + 63) break;
                    my_q = f >= e ? nq : my_q;
                    q = nq;
                }
                cur = q;
                ci = s_rlo[my_q] + (f - s_rex[my_q]);
                const int rglob = (int)rb + s_rid[my_q];
                ck = B.k0 + rglob / nj;
                cj = B.j0 + rglob % nj;
            };
            const float4 *fplane = a.factors + (size_t)B.layer * g.nx * g.ny * g.nz;
            int ni, nj_, nk_;
            locate(0, ni, nj_, nk_);
            uint32_t nv = (uint32_t)ni + (uint32_t)g.nx * ((uint32_t)nj_ + (uint32_t)g.ny * (uint32_t)nk_);
            float nw = 0.0f;
            float4 nfc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (MODE != INF_FILL || span_first <= pc) {
                nw = a.wet[nv];
                nfc = fplane[nv];
            }
            for (int f0 = 0; f0 < total; f0 += 64, pc++) {
                if (MODE == INF_FILL && pc > span_last) {
                    done = true;
                    break;
                }
                const int i = ni, j = nj_, k = nk_;
                const uint32_t v = nv;
                const bool inb = f0 + lane < total;
                const float w = nw * 10.0f;
                // factors of this (voxel, layer), written by the bounds pass
                const float4 fc = nfc;
                if (f0 + 64 < total) {
                    locate(f0 + 64, ni, nj_, nk_);
                    nv = (uint32_t)ni + (uint32_t)g.nx * ((uint32_t)nj_ + (uint32_t)g.ny * (uint32_t)nk_);
                    if (MODE != INF_FILL || span_first <= pc + 1) {
                        nw = a.wet[nv];
                        nfc = fplane[nv];
                    }
                }
                if (MODE == INF_FILL && pc < span_first) continue;
                const bool act = inb && !(w > L.cut);
                st_chunk++;
                st_visit += __popcll(__builtin_amdgcn_ballot_w64(inb));
                st_act += __popcll(__builtin_amdgcn_ballot_w64(act));
                if (!dh_wave_any(act)) continue;
                float px, py, pz, hx, hy, hz;
                dh_vox_xyz(g, i, j, k, px, py, pz);
                dh_img_to_head(g, px, py, pz, hx, hy, hz);
                const float cx = 1.0f - hz / a.vsadx, cy = 1.0f - hz / a.vsady;
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const float ux = fmaf(-sx[s], cx, hx), uy = fmaf(-sy[s], cy, hy);
                    const float r2 = fmaf(ux, ux, uy * uy);
                    const float dt = fmaf(tx[s], ux, ty[s] * uy);
                    const float d2 = fmaf(-dt, dt, r2);
                    const float ep = fmaf(fc.x, d2, fc.y), eh = fmaf(fc.z, d2, fc.w);
                    if (MODE == INF_PEAK) {
                        const float val = __builtin_amdgcn_exp2f(ep) + __builtin_amdgcn_exp2f(eh);
                        pk[s] = act ? fmaxf(pk[s], val) : pk[s];
                    } else {
                        // where both exponents are below log2(th) - 1.01 both
                        // terms are below th / 2 (v_exp_f32 errs by ~1 ulp),
                        // so the value is below th: a chunk with no other lane
                        // needs no exponentials and contributes nothing
                        if (!dh_wave_any(act && fmaxf(ep, eh) >= lth[s])) continue;
                        const float val = __builtin_amdgcn_exp2f(ep) + __builtin_amdgcn_exp2f(eh);
                        const bool keep = act && val >= th[s];
                        const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
                        if (MODE == INF_COUNT) {
                            cnt[s] += __popcll(m);
                            if (m) {
                                span_first = min(span_first, pc);
                                span_last = pc;
                            }
                        } else {
                            const PosT pos = cur64[s] + (PosT)__popcll(m & below);
                            if (keep && (P32 ? (uint32_t)pos < a.nnz32 : (int64_t)pos < lim[s])) {
                                a.col[pos] = v;
                                a.val[pos] = val;
                            }
                            cur64[s] += (PosT)__popcll(m);
                        }
                    }
                }
            }
            __syncthreads();
        }
        if (MODE == INF_COUNT && lane == 0 && a.spans)
            ((int2 *)a.spans)[item] = make_int2(span_first, span_last);
        if (a.stats && lane == 0) {
            atomicAdd(&a.stats[0], st_visit);
            atomicAdd(&a.stats[1], st_act);
            atomicAdd(&a.stats[2], st_chunk);
            atomicAdd(&a.stats[3], st_rows);
        }
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (sid[s] < 0) continue;
            if (MODE == INF_PEAK) {
                float m = pk[s];
                for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
                if (lane == 0 && m > 0.0f) atomicMax(&a.peak_bits[sid[s]], __float_as_uint(m));
            } else if (MODE == INF_COUNT) {
                if (lane == 0) a.counts[(size_t)sid[s] * a.C + chunk] = cnt[s];
            }
        }
    }
}

// Lower bound of each spot's peak: one wave per spot walks the spot axis
// through the volume in steps of half a voxel and evaluates, with the sweep's
// own formula, the voxel nearest to each axis point (active voxels only).
__global__ void __launch_bounds__(64)
k_axis_peak(dh_geom g, float vsadx, float vsady, int n_spots, const int *__restrict__ spot_layer,
            const dh_layer_k *__restrict__ layers, const float *__restrict__ spots,
            const float *__restrict__ wet, const float4 *__restrict__ factors, float hz0, float hz1,
            unsigned *__restrict__ peak_bits)
{
    const int q = blockIdx.x;
    const int lane = __lane_id();
    if (q >= n_spots) return;
    const int l = spot_layer[q];
    if (l < 0) return;
    const float sx = spots[8 * q], sy = spots[8 * q + 1], tx = spots[8 * q + 2], ty = spots[8 * q + 3];
    const float cut = layers[l].cut;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    const float step = 0.5f * g.s;
    const int nsteps = (int)ceilf((hz1 - hz0) / step) + 1;
    float m = 0.0f;
    for (int t = lane; t < nsteps; t += 64) {
        const float hza = hz0 + (float)t * step;
        float ax, ay, az;
        dh_head_to_img(g, sx * (1.0f - hza / vsadx), sy * (1.0f - hza / vsady), hza, ax, ay, az);
        const int i = (int)rintf((ax + g.iso[0]) * g.inv_s);
        const int j = (int)rintf((ay + g.iso[1]) * g.inv_s);
        const int k = (int)rintf((az + g.iso[2]) * g.inv_s);
        if (i < 0 || j < 0 || k < 0 || i >= g.nx || j >= g.ny || k >= g.nz) continue;
        const uint32_t v = (uint32_t)i + (uint32_t)g.nx * ((uint32_t)j + (uint32_t)g.ny * (uint32_t)k);
        if (wet[v] * 10.0f > cut) continue;
        float px, py, pz, hx, hy, hz;
        dh_vox_xyz(g, i, j, k, px, py, pz);
        dh_img_to_head(g, px, py, pz, hx, hy, hz);
        const float4 fc = factors[(size_t)l * nvox + v];
        const float cx = 1.0f - hz / vsadx, cy = 1.0f - hz / vsady;
        const float ux = fmaf(-sx, cx, hx), uy = fmaf(-sy, cy, hy);
        const float r2 = fmaf(ux, ux, uy * uy);
        const float dt = fmaf(tx, ux, ty * uy);
        const float d2 = fmaf(-dt, dt, r2);
        const float ep = fmaf(fc.x, d2, fc.y), eh = fmaf(fc.z, d2, fc.w);
        m = fmaxf(m, __builtin_amdgcn_exp2f(ep) + __builtin_amdgcn_exp2f(eh));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (lane == 0 && m > 0.0f) atomicMax(&peak_bits[q], __float_as_uint(m));
}

// Per layer and image row (k, j): first and last i with 10 * WET <= cut.
// One wave per row; the ballot of each 64-voxel chunk gives the extremes.
__global__ void __launch_bounds__(256)
k_active_rows(int nx, int ny, int nz, int n_layers, const dh_layer_k *__restrict__ layers,
              const float *__restrict__ wet, int2 *__restrict__ out)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = __lane_id();
    if (row >= ny * nz) return;
    const float *w = wet + (size_t)row * nx;
    for (int l = 0; l < n_layers; l++) {
        const float cut = layers[l].cut;
        int lo = nx, hi = -1;
        for (int i0 = 0; i0 < nx; i0 += 64) {
            const int i = i0 + lane;
            const bool act = i < nx && !(w[min(i, nx - 1)] * 10.0f > cut);
            const uint64_t m = __builtin_amdgcn_ballot_w64(act);
            if (m) {
                lo = min(lo, i0 + __ffsll((unsigned long long)m) - 1);
                hi = max(hi, i0 + 63 - __clzll((unsigned long long)m));
            }
        }
        if (lane == 0) out[(size_t)l * nz * ny + row] = make_int2(lo, hi);
    }
}

// Per layer, over the voxels the layer reaches (10 * WET <= 1.1 R80): the
// maxima of sigma_t, sigma_halo_total, P and H (float bits; all >= 0), which
// bound every spot's lateral dose profile in that layer, per workgroup (for
// the layer maxima) and per image row (row_bits [layer][nz * ny][4], zeroed
// by the caller: the per-row radius of the sweeps).  The folded factors are
// written for reached voxels only: the sweeps read a voxel's factors only
// where it is reached.  Workgroups with no reached voxel skip the layer.
__global__ void __launch_bounds__(256)
k_layer_bounds(dh_geom g, int n_layers, const dh_layer_k *__restrict__ layers,
               const float4 *__restrict__ luts, const float *__restrict__ wet,
               float *__restrict__ partial, float4 *__restrict__ factors, unsigned *__restrict__ row_bits)
{
    __shared__ float4 s_lut[DH_LUT_LENGTH];
    __shared__ float s_red[4][4];
    const uint32_t nvox = (uint32_t)g.nx * g.ny * g.nz;
    const uint32_t nrows = (uint32_t)g.ny * g.nz;
    const uint32_t v = blockIdx.x * 256 + threadIdx.x;
    const bool valid = v < nvox;
    const uint32_t vv = valid ? v : nvox - 1;
    const int i = (int)(vv % (uint32_t)g.nx);
    const uint32_t r = vv / (uint32_t)g.nx;
    const int j = (int)(r % (uint32_t)g.ny), k = (int)(r / (uint32_t)g.ny);
    const uint32_t r_first = __shfl(r, 0, 64), r_last = __shfl(r, 63, 64);
    float x, y, z;
    dh_vox_xyz(g, i, j, k, x, y, z);
    const float dxs = g.src[0] - x, dys = g.src[1] - y, dzs = g.src[2] - z;
    const float dist = sqrtf(dxs * dxs + dys * dys + dzs * dzs);
    const float w = wet[vv] * 10.0f;
    for (int l = 0; l < n_layers; l++) {
        const dh_layer_k L = layers[l];
        const bool act = valid && !(w > L.cut);
        // also the barrier before s_lut and s_red are rewritten
        if (!__syncthreads_or(act)) {
            if (threadIdx.x < 4) partial[((size_t)l * gridDim.x + blockIdx.x) * 4 + threadIdx.x] = 0.0f;
            continue;
        }
        for (int q = threadIdx.x; q < DH_LUT_LENGTH; q += 256) s_lut[q] = luts[(size_t)l * DH_LUT_LENGTH + q];
        __syncthreads();
        float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (act) {
            const dh_pb_factors f = dh_pencil_factors(s_lut, w, dist, L.r80, L.A, L.B, L.C);
            m[0] = fabsf(f.st); m[1] = fabsf(f.sht);
            m[2] = fmaxf(f.P, 0.0f); m[3] = fmaxf(f.H, 0.0f);
            factors[(size_t)l * nvox + v] =
                make_float4(f.ap, log2f(fmaxf(f.P, 0.0f)), f.ah, log2f(fmaxf(f.H, 0.0f)));
        }
        // per-row maxima: one atomic per quantity and wave when the wave lies
        // in one row, else per reached lane
        unsigned *rb = row_bits + (size_t)l * nrows * 4;
        if (r_first == r_last) {
#pragma unroll
            for (int q = 0; q < 4; q++) {
                float x2 = m[q];
                for (int off = 32; off > 0; off >>= 1) x2 = fmaxf(x2, __shfl_xor(x2, off, 64));
                if (__lane_id() == 0 && x2 > 0.0f) atomicMax(&rb[(size_t)r_first * 4 + q], __float_as_uint(x2));
            }
        } else if (act) {
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (m[q] > 0.0f) atomicMax(&rb[(size_t)r * 4 + q], __float_as_uint(m[q]));
        }
        // workgroup maxima -> one partial record per workgroup and layer
#pragma unroll
        for (int q = 0; q < 4; q++) {
            float x2 = m[q];
            for (int off = 32; off > 0; off >>= 1) x2 = fmaxf(x2, __shfl_xor(x2, off, 64));
            if (__lane_id() == 0) s_red[threadIdx.x >> 6][q] = x2;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int q = threadIdx.x;
            const float x2 = fmaxf(fmaxf(s_red[0][q], s_red[1][q]), fmaxf(s_red[2][q], s_red[3][q]));
            partial[((size_t)l * gridDim.x + blockIdx.x) * 4 + q] = x2;
        }
    }
}

// Reduces the partial maxima: one workgroup per (layer, quantity).
__global__ void __launch_bounds__(256)
k_reduce_bounds(const float *__restrict__ partial, int nblocks, float *__restrict__ out)
{
    const int l = blockIdx.x >> 2, q = blockIdx.x & 3;
    __shared__ float s_m[4];
    float m = 0.0f;
    for (int b = threadIdx.x; b < nblocks; b += 256) m = fmaxf(m, partial[((size_t)l * nblocks + b) * 4 + q]);
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (__lane_id() == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[4 * l + q] = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
}

__global__ void k_zero_words(unsigned *p, size_t n)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = 0u;
}

__global__ void k_fill_words(unsigned *p, size_t n, unsigned v)
{
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) p[t] = v;
}

// Per spot, on the device (no host round trip between the passes): the
// sweep radius and threshold from the spot's peak (bits in peak_bits) and its
// layer's maxima bnd[4 l + {0..3}] = {sigma_t, sigma_ht, P, H}.
// mode 0 (PEAK sweep; peak_bits = the axis lower bound pa): threshold pa;
// beyond the radius both Gaussian terms of every voxel are below pa / 2, so
// no voxel there can exceed pa (3 sigma_ht + s when the axis met no voxel).
// mode 1 (COUNT / FILL; peak_bits = the exact peak): threshold rel * peak;
// beyond the radius both terms are below thr / 2; the layer's smallest
// threshold goes to lthr_bits (float bits, atomicMin; thresholds > 0).
__global__ void __launch_bounds__(256)
k_spot_radius(int n, int mode, const int *__restrict__ spot_layer, const unsigned *__restrict__ peak_bits,
              const float *__restrict__ bnd, float rel, float s, float *__restrict__ thr, float *__restrict__ radius,
              unsigned *__restrict__ lthr_bits)
{
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const int l = spot_layer[q];
    const float v = __uint_as_float(peak_bits[q]);
    if (mode == 0) {
        thr[q] = v;
        if (l < 0) {
            radius[q] = 0.0f;
            return;
        }
        const float st = bnd[4 * l], sht = bnd[4 * l + 1], Pm = bnd[4 * l + 2], Hm = bnd[4 * l + 3];
        if (!(v > 0.0f)) {
            radius[q] = 3.0f * sht + s;
            return;
        }
        const float rp = Pm > 0.5f * v ? st * sqrtf(2.0f * logf(2.0f * Pm / v)) : 0.0f;
        const float rh = Hm > 0.5f * v ? sht * sqrtf(2.0f * logf(2.0f * Hm / v)) : 0.0f;
        radius[q] = 1.001f * fmaxf(rp, rh) + s;
    } else {
        if (l < 0 || !(v > 0.0f)) {
            thr[q] = INFINITY;
            radius[q] = 0.0f;
            return;
        }
        const float t = rel * v;
        thr[q] = t;
        const float st = bnd[4 * l], sht = bnd[4 * l + 1], Pm = bnd[4 * l + 2], Hm = bnd[4 * l + 3];
        const float rp = Pm > 0.5f * t ? st * sqrtf(2.0f * logf(2.0f * Pm / t)) : 0.0f;
        const float rh = Hm > 0.5f * t ? sht * sqrtf(2.0f * logf(2.0f * Hm / t)) : 0.0f;
        radius[q] = fmaxf(rp, rh) + s;
        atomicMin(&lthr_bits[l], __float_as_uint(t));
    }
}

// The sweeps' work blocks from the spots' radii, on the device: block b's
// lateral box (its spot axes over the volume's head depths, widened by the
// largest radius) and the image rows the box's prism covers; written to
// position pos[b] of the block list (the order the persistent queue hands
// them out in).
struct blk_consts {
    float cmin_x, cmax_x, cmin_y, cmax_y, hz_min, hz_max;
};

__global__ void __launch_bounds__(256)
k_make_blocks(dh_geom g, int nb, int S, blk_consts k, const int *__restrict__ block_layer,
              const int *__restrict__ block_n, const int *__restrict__ block_spots, const float *__restrict__ spots,
              const float *__restrict__ radius, const int *__restrict__ pos, dh_infl_block *__restrict__ out)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nb) return;
    dh_infl_block B;
    B.layer = block_layer[b];
    B.first = b;
    B.n = block_n[b];
    B.pad_ = 0;
    float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
    float e0 = 1e30f, e1 = -1e30f, f0 = 1e30f, f1 = -1e30f, Rb = 0.0f;
    for (int q = 0; q < S; q++) {
        const int id = block_spots[b * S + q];
        if (id < 0) continue;
        const float sx = spots[8 * id], sy = spots[8 * id + 1];
        const float R = radius[id] * 1.002f + 0.5f * g.s;
        const float ax0 = fminf(sx * k.cmin_x, sx * k.cmax_x), ax1 = fmaxf(sx * k.cmin_x, sx * k.cmax_x);
        const float ay0 = fminf(sy * k.cmin_y, sy * k.cmax_y), ay1 = fmaxf(sy * k.cmin_y, sy * k.cmax_y);
        x0 = fminf(x0, ax0 - R); x1 = fmaxf(x1, ax1 + R);
        y0 = fminf(y0, ay0 - R); y1 = fmaxf(y1, ay1 + R);
        e0 = fminf(e0, ax0); e1 = fmaxf(e1, ax1);
        f0 = fminf(f0, ay0); f1 = fmaxf(f1, ay1);
        Rb = fmaxf(Rb, R);
    }
    B.bx0 = x0; B.bx1 = x1; B.by0 = y0; B.by1 = y1;
    B.ex0 = e0; B.ex1 = e1; B.ey0 = f0; B.ey1 = f1; B.R = Rb;
    B.pad1_ = B.pad2_ = B.pad3_ = 0.0f;
    float kmin = 1e30f, kmax = -1e30f, jmin = 1e30f, jmax = -1e30f;
    for (int c = 0; c < 8; c++) {
        float px, py, pz;
        dh_head_to_img(g, (c & 1) ? x1 : x0, (c & 2) ? y1 : y0, (c & 4) ? k.hz_max : k.hz_min, px, py, pz);
        const float fj = (py + g.iso[1]) / g.s, fk = (pz + g.iso[2]) / g.s;
        jmin = fminf(jmin, fj); jmax = fmaxf(jmax, fj);
        kmin = fminf(kmin, fk); kmax = fmaxf(kmax, fk);
    }
    B.j0 = max(0, (int)floorf(jmin) - 1); B.j1 = min(g.ny - 1, (int)ceilf(jmax) + 1);
    B.k0 = max(0, (int)floorf(kmin) - 1); B.k1 = min(g.nz - 1, (int)ceilf(kmax) + 1);
    if (B.j0 > B.j1 || B.k0 > B.k1) { B.j0 = 0; B.j1 = -1; B.k0 = 0; B.k1 = 0; }
    out[pos[b]] = B;
}

// Smallest box in the head frame: min/max over the volume corners of the
// head coordinates (for the divergence scaling of the spot axes).
static void head_depth_range(const dh_geom &g, float &hz_min, float &hz_max)
{
    hz_min = 1e30f; hz_max = -1e30f;
    for (int c = 0; c < 8; c++) {
        float x, y, z, hx, hy, hz;
        dh_vox_xyz(g, (c & 1) ? g.nx - 1 : 0, (c & 2) ? g.ny - 1 : 0, (c & 4) ? g.nz - 1 : 0, x, y, z);
        dh_img_to_head(g, x, y, z, hx, hy, hz);
        hz_min = std::min(hz_min, hz);
        hz_max = std::max(hz_max, hz);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct dh_influence {
    int dev = 0;
    dh_geom g;
    float vsadx = 0, vsady = 0;
    int S = 16, C = 8;
    int cus = 0;
    int n_spots = 0;
    float rel = 1e-4f;
    std::vector<dh_layer_k> layers;
    std::vector<float4> luts;
    std::vector<float> spots;        // 8 floats per sorted spot
    std::vector<int> spot_layer;
    std::vector<float> sig_max, dose_max;   // per-layer bounds
    std::vector<std::vector<float4>> cells;  // per layer: depth-resolved table bounds
    std::vector<float> lthr;         // per layer: smallest threshold of the last count (inf: none yet)
    std::vector<int> block_layer, block_first, block_n;
    std::vector<int> block_spots;    // S per block
    dh_layer_k *d_layers = nullptr;
    float4 *d_luts = nullptr;
    float *d_spots = nullptr;
    int *d_block_spots = nullptr;
    dh_infl_block *d_blocks = nullptr;
    float *d_thr = nullptr;
    unsigned *d_peak = nullptr;
    int *d_counts = nullptr;
    int *d_spans = nullptr;
    int64_t *d_offsets = nullptr;
    int *d_queue = nullptr;
    int2 *d_active = nullptr;
    int *d_spot_layer = nullptr;
    unsigned long long *d_stats = nullptr;   // 4 counters per pass
    float4 *d_factors = nullptr;             // [layer][voxel]
    unsigned *d_row_bits = nullptr;          // [layer][row][4] per-row maxima (float bits)
    float *d_bnd = nullptr;                  // [layer][4] maxima of sigma_t, sigma_ht, P, H
    float *d_radius = nullptr;               // per spot: sweep radius of the last pass set up
    unsigned *d_lthr = nullptr;              // per layer: smallest threshold (float bits)
    int *d_block_layer = nullptr, *d_block_n = nullptr;
    int *d_block_pos = nullptr;              // per block: its position in the queue order
    std::vector<float> peak;         // host copy after the peak pass
    std::vector<float> radius;       // truncation radius per spot (COUNT = FILL)
    std::vector<int> order;          // caller's spot row of each matrix row
    std::vector<int64_t> offsets;    // host copy of the exclusive scan
    int64_t nnz = -1;
    int fill_blocks = -1;            // blocks of the prepared fill pass in d_blocks (-1: none)
    bool counted_blocks = false;     // d_blocks holds the COUNT pass's blocks
    // footprint-limited WET (dh_influence_wet): the inner head-frame box
    // where the WET is exact, the RLSP it came from and the buffer it went to
    bool fp_on = false;
    const float *fp_rlsp = nullptr;
    const float *fp_wet = nullptr;
    dh_hbox fp_inner{0.0f, 0.0f, 0.0f, 0.0f, 0};
    int64_t fp_fallbacks = 0;
    // pinned host copies of the count's results (hipHostMalloc on first use)
    int *h_cnt = nullptr;
    unsigned *h_pk = nullptr;
    float *h_thr = nullptr, *h_r = nullptr;
};

static std::vector<dh_infl_block> make_blocks(const dh_influence *f, const std::vector<float> &r);

dh_influence *dh_influence_make(int dev, const dh_geom &g, float vsadx, float vsady,
                                const std::vector<dh_layer_k> &layers, const std::vector<float4> &luts,
                                const std::vector<float> &spots, int n_spots, int S, int C, float rel)
{
    DH_REQUIRE(S == 4 || S == 8 || S == 16, "spots per block must be 4, 8 or 16");
    DH_REQUIRE(C >= 1 && C <= 256, "row chunks per block must be 1..256");
    dh_influence *f = new dh_influence();
    f->dev = dev; f->g = g; f->vsadx = vsadx; f->vsady = vsady; f->S = S; f->C = C;
    f->n_spots = n_spots; f->rel = rel;
    f->layers = layers; f->luts = luts; f->spots = spots;
    // distances from the source to the volume (bounding sphere)
    float dmin = 1e30f, dmax = 0.0f;
    for (int c = 0; c < 8; c++) {
        float x, y, z;
        dh_vox_xyz(g, (c & 1) ? g.nx - 1 : 0, (c & 2) ? g.ny - 1 : 0, (c & 4) ? g.nz - 1 : 0, x, y, z);
        const float d = std::sqrt((g.src[0] - x) * (g.src[0] - x) + (g.src[1] - y) * (g.src[1] - y) +
                                  (g.src[2] - z) * (g.src[2] - z));
        dmin = std::min(dmin, d);
        dmax = std::max(dmax, d);
    }
    const float diag = 0.5f * g.s * std::sqrt((float)(g.nx * g.nx + g.ny * g.ny + g.nz * g.nz));
    dmin = std::max(0.0f, dmin - diag);     // the source may face any point of the box
    f->spot_layer.assign(n_spots, -1);
    for (size_t l = 0; l < layers.size(); l++) {
        float sm, dm;
        dh_infl_layer_bounds(layers[l], &luts[l * DH_LUT_LENGTH], dmin, dmax, sm, dm);
        f->sig_max.push_back(sm);
        f->dose_max.push_back(dm);
        f->cells.push_back(dh_infl_depth_cells(layers[l], &luts[l * DH_LUT_LENGTH], dmin, dmax));
        f->lthr.push_back(INFINITY);
        // blocks of S spatially adjacent spots: sort the layer's spots by a
        // Morton key of their position on a 4 mm grid
        const int a0 = layers[l].start, n = layers[l].n;
        std::vector<std::pair<uint32_t, int>> key(n);
        for (int q = 0; q < n; q++) {
            const int id = a0 + q;
            f->spot_layer[id] = (int)l;
            auto qz = [](float v) {
                int c = (int)floorf(v / 4.0f) + 512;
                return (uint32_t)std::min(std::max(c, 0), 1023);
            };
            uint32_t xq = qz(spots[8 * id]), yq = qz(spots[8 * id + 1]), m = 0;
            for (int b = 0; b < 10; b++) m |= (((xq >> b) & 1u) << (2 * b)) | (((yq >> b) & 1u) << (2 * b + 1));
            key[q] = {m, id};
        }
        std::stable_sort(key.begin(), key.end(), [](auto &x, auto &y) { return x.first < y.first; });
        for (int q = 0; q < n; q += S) {
            f->block_layer.push_back((int)l);
            f->block_first.push_back(q);
            f->block_n.push_back(std::min(S, n - q));
            for (int s = 0; s < S; s++) f->block_spots.push_back(q + s < n ? key[q + s].second : -1);
        }
    }
    const size_t nb = f->block_layer.size();
    try {
        DH_HIP(hipSetDevice(dev));
        if (!layers.empty()) {
            DH_HIP(hipMalloc(&f->d_layers, sizeof(dh_layer_k) * layers.size()));
            DH_HIP(hipMalloc(&f->d_luts, sizeof(float4) * luts.size()));
            DH_HIP(hipMemcpy(f->d_layers, layers.data(), sizeof(dh_layer_k) * layers.size(), hipMemcpyHostToDevice));
            DH_HIP(hipMemcpy(f->d_luts, luts.data(), sizeof(float4) * luts.size(), hipMemcpyHostToDevice));
        }
        DH_HIP(hipMalloc(&f->d_spots, sizeof(float) * std::max<size_t>(spots.size(), 8)));
        if (!spots.empty())
            DH_HIP(hipMemcpy(f->d_spots, spots.data(), sizeof(float) * spots.size(), hipMemcpyHostToDevice));
        DH_HIP(hipMalloc(&f->d_block_spots, sizeof(int) * std::max<size_t>(f->block_spots.size(), 1)));
        if (nb)
            DH_HIP(hipMemcpy(f->d_block_spots, f->block_spots.data(), sizeof(int) * f->block_spots.size(), hipMemcpyHostToDevice));
        DH_HIP(hipMalloc(&f->d_blocks, sizeof(dh_infl_block) * std::max<size_t>(nb, 1)));
        DH_HIP(hipMalloc(&f->d_thr, sizeof(float) * std::max(n_spots, 1)));
        DH_HIP(hipMalloc(&f->d_peak, sizeof(unsigned) * std::max(n_spots, 1)));
        DH_HIP(hipMalloc(&f->d_counts, sizeof(int) * std::max<size_t>((size_t)n_spots * C, 1)));
        DH_HIP(hipMalloc(&f->d_spans, sizeof(int2) * std::max<size_t>(nb * (size_t)C, 1)));
        DH_HIP(hipMalloc(&f->d_offsets, sizeof(int64_t) * ((size_t)n_spots * C + 1)));
        DH_HIP(hipMalloc(&f->d_queue, 256));
        DH_HIP(hipMalloc(&f->d_stats, sizeof(unsigned long long) * 12));
        DH_HIP(hipMalloc(&f->d_factors, sizeof(float4) * std::max<size_t>(layers.size(), 1) *
                                            (size_t)g.nx * g.ny * g.nz));
        DH_HIP(hipMalloc(&f->d_active, sizeof(int2) * std::max<size_t>(layers.size(), 1) * g.ny * g.nz));
        DH_HIP(hipMalloc(&f->d_row_bits, sizeof(unsigned) * 4 * std::max<size_t>(layers.size(), 1) * g.ny * g.nz));
        DH_HIP(hipMalloc(&f->d_bnd, sizeof(float) * 4 * std::max<size_t>(layers.size(), 1)));
        DH_HIP(hipMalloc(&f->d_radius, sizeof(float) * std::max(n_spots, 1)));
        DH_HIP(hipMalloc(&f->d_lthr, sizeof(unsigned) * std::max<size_t>(layers.size(), 1)));
        DH_HIP(hipMalloc(&f->d_block_layer, sizeof(int) * std::max<size_t>(nb, 1)));
        DH_HIP(hipMalloc(&f->d_block_n, sizeof(int) * std::max<size_t>(nb, 1)));
        DH_HIP(hipMalloc(&f->d_block_pos, sizeof(int) * std::max<size_t>(nb, 1)));
        if (nb) {
            DH_HIP(hipMemcpy(f->d_block_layer, f->block_layer.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
            DH_HIP(hipMemcpy(f->d_block_n, f->block_n.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
            // the queue order: largest estimated work first (block boxes at a
            // radius of 3 sigma_max of the layer); every pass keeps this order
            std::vector<float> r_est(n_spots, 0.0f);
            for (int q = 0; q < n_spots; q++)
                if (f->spot_layer[q] >= 0) r_est[q] = 3.0f * f->sig_max[f->spot_layer[q]] + g.s;
            const std::vector<dh_infl_block> est = make_blocks(f, r_est);
            std::vector<int> pos(nb);
            for (size_t i = 0; i < nb; i++) pos[est[i].first] = (int)i;
            DH_HIP(hipMemcpy(f->d_block_pos, pos.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
        }
        DH_HIP(hipMalloc(&f->d_spot_layer, sizeof(int) * std::max(n_spots, 1)));
        if (n_spots)
            DH_HIP(hipMemcpy(f->d_spot_layer, f->spot_layer.data(), sizeof(int) * n_spots, hipMemcpyHostToDevice));
    } catch (...) {
        void dh_influence_free(dh_influence *);
        dh_influence_free(f);
        throw;
    }
    return f;
}

void dh_influence_free(dh_influence *f)
{
    if (!f) return;
    (void)hipSetDevice(f->dev);
    void *ps[] = {f->d_layers, f->d_luts, f->d_spots, f->d_block_spots, f->d_blocks, f->d_thr,
                  f->d_peak, f->d_counts, f->d_spans, f->d_offsets, f->d_queue, f->d_active, f->d_stats,
                  f->d_factors, f->d_spot_layer, f->d_row_bits, f->d_bnd, f->d_radius, f->d_lthr,
                  f->d_block_layer, f->d_block_n, f->d_block_pos};
    for (void *p : ps)
        if (p) (void)hipFree(p);
    void *hs[] = {f->h_cnt, f->h_pk, f->h_thr, f->h_r};
    for (void *p : hs)
        if (p) (void)hipHostFree(p);
    delete f;
}

// Blocks' boxes for a pass: radius per spot r[s] (mm).
static std::vector<dh_infl_block> make_blocks(const dh_influence *f, const std::vector<float> &r)
{
    const dh_geom &g = f->g;
    float hz_min, hz_max;
    head_depth_range(g, hz_min, hz_max);
    const float cmin_x = std::min(1.0f - hz_max / f->vsadx, 1.0f - hz_min / f->vsadx);
    const float cmax_x = std::max(1.0f - hz_max / f->vsadx, 1.0f - hz_min / f->vsadx);
    const float cmin_y = std::min(1.0f - hz_max / f->vsady, 1.0f - hz_min / f->vsady);
    const float cmax_y = std::max(1.0f - hz_max / f->vsady, 1.0f - hz_min / f->vsady);
    std::vector<dh_infl_block> out(f->block_layer.size());
    for (size_t b = 0; b < out.size(); b++) {
        dh_infl_block &B = out[b];
        B.layer = f->block_layer[b]; B.first = (int)b; B.n = f->block_n[b]; B.pad_ = 0;
        float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
        float e0 = 1e30f, e1 = -1e30f, f0 = 1e30f, f1 = -1e30f, Rb = 0.0f;
        for (int s = 0; s < f->S; s++) {
            const int id = f->block_spots[b * f->S + s];
            if (id < 0) continue;
            const float sx = f->spots[8 * id], sy = f->spots[8 * id + 1];
            const float R = r[id] * 1.002f + 0.5f * g.s;
            x0 = std::min(x0, std::min(sx * cmin_x, sx * cmax_x) - R);
            x1 = std::max(x1, std::max(sx * cmin_x, sx * cmax_x) + R);
            y0 = std::min(y0, std::min(sy * cmin_y, sy * cmax_y) - R);
            y1 = std::max(y1, std::max(sy * cmin_y, sy * cmax_y) + R);
            e0 = std::min(e0, std::min(sx * cmin_x, sx * cmax_x));
            e1 = std::max(e1, std::max(sx * cmin_x, sx * cmax_x));
            f0 = std::min(f0, std::min(sy * cmin_y, sy * cmax_y));
            f1 = std::max(f1, std::max(sy * cmin_y, sy * cmax_y));
            Rb = std::max(Rb, R);
        }
        B.bx0 = x0; B.bx1 = x1; B.by0 = y0; B.by1 = y1;
        B.ex0 = e0; B.ex1 = e1; B.ey0 = f0; B.ey1 = f1; B.R = Rb;
        B.pad1_ = B.pad2_ = B.pad3_ = 0.0f;
        // image-space bounding rows of the prism box x [hz_min, hz_max]
        float kmin = 1e30f, kmax = -1e30f, jmin = 1e30f, jmax = -1e30f;
        for (int c = 0; c < 8; c++) {
            float px, py, pz;
            dh_head_to_img(g, (c & 1) ? x1 : x0, (c & 2) ? y1 : y0, (c & 4) ? hz_max : hz_min, px, py, pz);
            const float fj = (py + g.iso[1]) / g.s, fk = (pz + g.iso[2]) / g.s;
            jmin = std::min(jmin, fj); jmax = std::max(jmax, fj);
            kmin = std::min(kmin, fk); kmax = std::max(kmax, fk);
        }
        B.j0 = std::max(0, (int)floorf(jmin) - 1); B.j1 = std::min(g.ny - 1, (int)ceilf(jmax) + 1);
        B.k0 = std::max(0, (int)floorf(kmin) - 1); B.k1 = std::min(g.nz - 1, (int)ceilf(kmax) + 1);
        if (B.j0 > B.j1 || B.k0 > B.k1) { B.j0 = 0; B.j1 = -1; B.k0 = 0; B.k1 = 0; }
    }
    // the persistent queue hands out the most expensive blocks first
    std::stable_sort(out.begin(), out.end(), [](const dh_infl_block &p, const dh_infl_block &q) {
        const int64_t cp = (int64_t)(p.j1 - p.j0 + 1) * (p.k1 - p.k0 + 1) * (int64_t)(p.bx1 - p.bx0 + p.by1 - p.by0);
        const int64_t cq = (int64_t)(q.j1 - q.j0 + 1) * (q.k1 - q.k0 + 1) * (int64_t)(q.bx1 - q.bx0 + q.by1 - q.by0);
        return cp > cq;
    });
    return out;
}

// Kernel launches of one pass over the blocks already in f->d_blocks (no
// host copies: safe to enqueue under stream capture).
template <int MODE>
static void launch_kernels(dh_influence *f, int n_blocks, const float *d_wet, uint32_t *d_col, float *d_val,
                           hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, (unsigned *)f->d_queue, (size_t)1);
    dh_infl_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = f->g; a.vsadx = f->vsadx; a.vsady = f->vsady;
    a.n_items = n_blocks * f->C; a.C = f->C;
    a.blocks = f->d_blocks; a.block_spots = f->d_block_spots; a.layers = f->d_layers;
    a.luts = f->d_luts; a.spots = f->d_spots; a.thr = f->d_thr; a.wet = d_wet;
    a.peak_bits = f->d_peak; a.counts = f->d_counts; a.spans = f->d_spans; a.offsets = f->d_offsets;
    a.active = f->d_active;
    a.factors = f->d_factors;
    a.row_bounds = (const float4 *)f->d_row_bits;
    a.stats = f->d_stats + 4 * MODE;
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, s, (unsigned *)a.stats, (size_t)8);
    a.col = d_col; a.val = d_val; a.queue = f->d_queue;
    if (a.n_items == 0) return;
    if (!f->cus) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, f->dev) == hipSuccess && v > 0)
            f->cus = v;
        else
            f->cus = 256;
    }
    const int cus = f->cus;
    const unsigned grid = (unsigned)std::min(a.n_items, cus * 32);
    const char *name = MODE == INF_PEAK ? "influence_peak" : MODE == INF_COUNT ? "influence_count" : "influence_fill";
    dh_timed t(name, s);
    // the fill takes 32-bit positions when the matrix's last slot (absolute,
    // counting the `base` of a CSR shared between fields) lies below 2^32;
    // stores are bounded by that end
    const int64_t end = f->offsets.empty() ? 0 : f->offsets.back();
    const bool p32 = MODE == INF_FILL && f->nnz >= 0 && end < ((int64_t)1 << 32);
    a.nnz32 = p32 ? (uint32_t)end : 0u;
    if constexpr (MODE == INF_FILL) {
        if (p32) {
            if (f->S == 4) hipLaunchKernelGGL((k_influence<4, MODE, true>), dim3(grid), dim3(64), 0, s, a);
            else if (f->S == 8) hipLaunchKernelGGL((k_influence<8, MODE, true>), dim3(grid), dim3(64), 0, s, a);
            else hipLaunchKernelGGL((k_influence<16, MODE, true>), dim3(grid), dim3(64), 0, s, a);
            DH_HIP(hipGetLastError());
            return;
        }
    }
    if (f->S == 4) hipLaunchKernelGGL((k_influence<4, MODE>), dim3(grid), dim3(64), 0, s, a);
    else if (f->S == 8) hipLaunchKernelGGL((k_influence<8, MODE>), dim3(grid), dim3(64), 0, s, a);
    else hipLaunchKernelGGL((k_influence<16, MODE>), dim3(grid), dim3(64), 0, s, a);
    DH_HIP(hipGetLastError());
}

// The work blocks of the next pass from the spots' radii in d_radius, on
// the device (stream-ordered: no host copy, no synchronisation).
static void device_blocks(dh_influence *f, hipStream_t s)
{
    const int nb = (int)f->block_layer.size();
    f->fill_blocks = -1;
    f->counted_blocks = false;
    if (!nb) return;
    float hz_min, hz_max;
    head_depth_range(f->g, hz_min, hz_max);
    blk_consts k;
    k.cmin_x = std::min(1.0f - hz_max / f->vsadx, 1.0f - hz_min / f->vsadx);
    k.cmax_x = std::max(1.0f - hz_max / f->vsadx, 1.0f - hz_min / f->vsadx);
    k.cmin_y = std::min(1.0f - hz_max / f->vsady, 1.0f - hz_min / f->vsady);
    k.cmax_y = std::max(1.0f - hz_max / f->vsady, 1.0f - hz_min / f->vsady);
    k.hz_min = hz_min;
    k.hz_max = hz_max;
    hipLaunchKernelGGL(k_make_blocks, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, f->g, nb, f->S, k,
                       f->d_block_layer, f->d_block_n, f->d_block_spots, f->d_spots, f->d_radius, f->d_block_pos,
                       f->d_blocks);
    DH_HIP(hipGetLastError());
}

// Per-spot radii and thresholds of a pass on the device (k_spot_radius).
static void device_radius(dh_influence *f, int mode, hipStream_t s)
{
    const int n = f->n_spots;
    if (!n) return;
    if (mode == 1) {
        const size_t nl = std::max<size_t>(f->layers.size(), 1);
        hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(256), 0, s, f->d_lthr, nl, 0x7f800000u);
    }
    hipLaunchKernelGGL(k_spot_radius, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, mode, f->d_spot_layer,
                       (const unsigned *)f->d_peak, (const float *)f->d_bnd, f->rel, f->g.s, f->d_thr, f->d_radius,
                       f->d_lthr);
    DH_HIP(hipGetLastError());
}

// Radius beyond which every voxel of layer l, whatever its WET, has a dose
// below thr: dose <= (P + H) e^(-d^2 / 2 sigma^2) with the layer's table
// bounds P + H <= dose_max and sigma_t, sigma_ht <= sig_max.
static float table_radius(const dh_influence *f, int l, float thr)
{
    return dh_infl_depth_radius(f->cells[l], thr);
}

// Divergence factors 1 - hz / VSAD over the volume's head depths.
static void divergence_range(const dh_influence *f, float &cx0, float &cx1, float &cy0, float &cy1)
{
    float hz_min, hz_max;
    head_depth_range(f->g, hz_min, hz_max);
    cx0 = std::min(1.0f - hz_max / f->vsadx, 1.0f - hz_min / f->vsadx);
    cx1 = std::max(1.0f - hz_max / f->vsadx, 1.0f - hz_min / f->vsadx);
    cy0 = std::min(1.0f - hz_max / f->vsady, 1.0f - hz_min / f->vsady);
    cy1 = std::max(1.0f - hz_max / f->vsady, 1.0f - hz_min / f->vsady);
}

// Smoothed WET of the field on the device: ray trace + lateral smoothing
// (CudaClasses.cu:180-228, IMPTClasses.cu:139-202) limited to a head-frame
// box, or on every voxel when box is null.  The ray trace covers the box
// widened by the smoothing's reach (9 mm radius, rounded to the nearest
// voxel), so every value inside the box equals the full-grid one bit for bit.
static void field_wet(dh_influence *f, const float *d_rlsp, float *d_wet, const dh_hbox *box, hipStream_t s)
{
    const dh_geom &g = f->g;
    const dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
    float4 *vpad = (float4 *)dh_scratch(f->dev, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
    float *raw = (float *)dh_scratch(f->dev, DH_SLOT_WET_RAW, sizeof(float) * (size_t)g.nx * g.ny * g.nz);
    dh_launch_pad_quad(d_rlsp, vpad, g.nx, g.ny, g.nz, 0.0f, s);      // IMPTClasses.cu:292, border 0
    if (box) {
        // bricks of 16 x 4 x 4 voxels: a wave covers a 16 x 4 patch, so few
        // waves straddle the box edge (the best measured tile for the
        // (160, 256, 256) fields; "raytrace_box" / "smooth_box" override it)
        const std::string key = dh_tune_key_geom(g);
        const float reach = 10.0f + g.s;
        const dh_hbox outer{box->x0 - reach, box->x1 + reach, box->y0 - reach, box->y1 + reach, 1};
        dh_launch_raytrace(g, vpad, raw, s, dh_tune_lookup("raytrace_box", key, 5), &outer);
        dh_launch_smooth(g, raw, d_wet, s, dh_tune_lookup("smooth_box", key, 5), box);
    } else {
        dh_launch_raytrace(g, vpad, raw, s);
        dh_launch_smooth(g, raw, d_wet, s);
    }
}

// Footprint of the field: the spot axes widened per layer by the table
// radius at an estimated threshold, rel * 0.05 * (a lower bound of the
// primary dose at the entrance, IDD(0) / (2 pi sig_max^2) * (1 - 0.9)),
// scaled by margin_scale.  dh_influence_count checks every spot's table
// radius at its actual threshold against this box and recomputes the full
// WET when one does not fit, so the matrix never depends on the estimate.
static void footprint_wet(dh_influence *f, const float *d_rlsp, float margin_scale, float *d_wet, hipStream_t s)
{
    const dh_geom &g = f->g;
    float cx0, cx1, cy0, cy1;
    divergence_range(f, cx0, cx1, cy0, cy1);
    float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
    for (size_t l = 0; l < f->layers.size(); l++) {
        // threshold: the layer's smallest one of the last count, else an
        // estimate from the entrance dose
        float thr = f->lthr[l];
        if (!std::isfinite(thr)) {
            const float idd0 = f->luts[l * DH_LUT_LENGTH].z;
            const float sm = f->sig_max[l];
            thr = f->rel * 0.05f * 0.1f * std::max(idd0, 0.0f) / (2.0f * DH_PI_F * sm * sm);
        }
        const float R = table_radius(f, (int)l, thr);
        const float M = margin_scale * (std::isfinite(R) ? R : 1e4f) + 0.87f * g.s;
        const dh_layer_k &L = f->layers[l];
        for (int q = L.start; q < L.start + L.n; q++) {
            const float sx = f->spots[8 * q], sy = f->spots[8 * q + 1];
            x0 = std::min(x0, std::min(sx * cx0, sx * cx1) - M);
            x1 = std::max(x1, std::max(sx * cx0, sx * cx1) + M);
            y0 = std::min(y0, std::min(sy * cy0, sy * cy1) - M);
            y1 = std::max(y1, std::max(sy * cy0, sy * cy1) + M);
        }
    }
    if (x0 > x1) { x0 = 1.0f; x1 = -1.0f; y0 = 1.0f; y1 = -1.0f; }   // no spots: nothing inside
    const dh_hbox inner{x0, x1, y0, y1, 1};
    field_wet(f, d_rlsp, d_wet, &inner, s);
    f->fp_on = true;
    f->fp_rlsp = d_rlsp;
    f->fp_wet = d_wet;
    f->fp_inner = inner;
}

// True when every spot's table radius at its threshold fits the footprint:
// then every voxel outside it is below every threshold (and below every
// peak), so the matrix equals the one of the full-grid WET.
static bool footprint_holds(const dh_influence *f, const std::vector<float> &thr)
{
    float cx0, cx1, cy0, cy1;
    divergence_range(f, cx0, cx1, cy0, cy1);
    const dh_hbox &b = f->fp_inner;
    // one radius per layer, at the layer's smallest threshold (the largest radius)
    std::vector<float> lmin(f->layers.size(), INFINITY), Rl(f->layers.size(), 0.0f);
    for (int q = 0; q < f->n_spots; q++)
        if (f->spot_layer[q] >= 0) lmin[f->spot_layer[q]] = std::min(lmin[f->spot_layer[q]], thr[q]);
    for (size_t l = 0; l < lmin.size(); l++)
        Rl[l] = std::isfinite(lmin[l]) ? table_radius(f, (int)l, lmin[l]) + 0.87f * f->g.s : 0.0f;
    for (int q = 0; q < f->n_spots; q++) {
        const int l = f->spot_layer[q];
        if (l < 0 || !std::isfinite(thr[q])) continue;
        const float R = Rl[l];
        const float sx = f->spots[8 * q], sy = f->spots[8 * q + 1];
        if (!(std::min(sx * cx0, sx * cx1) - R >= b.x0 && std::max(sx * cx0, sx * cx1) + R <= b.x1 &&
              std::min(sy * cy0, sy * cy1) - R >= b.y0 && std::max(sy * cy0, sy * cy1) + R <= b.y1))
            return false;
    }
    return true;
}

// Per layer and image row, the x-interval of voxels the layer reaches.
static void active_rows(dh_influence *f, const float *d_wet, hipStream_t s)
{
    const int nrow = f->g.ny * f->g.nz;
    if (f->layers.empty()) return;
    dh_timed t("influence_rows", s);
    hipLaunchKernelGGL(k_active_rows, dim3((nrow + 3) / 4), dim3(256), 0, s, f->g.nx, f->g.ny, f->g.nz,
                       (int)f->layers.size(), f->d_layers, d_wet, f->d_active);
    DH_HIP(hipGetLastError());
}

// Per-layer maxima of sigma_t, sigma_ht, P, H over the voxels the layer
// reaches, from the actual WET.
static void wet_bounds(dh_influence *f, const float *d_wet, hipStream_t s)
{
    const size_t nl = f->layers.size();
    if (!nl) return;
    const size_t nvox = (size_t)f->g.nx * f->g.ny * f->g.nz;
    const unsigned nblk = (unsigned)((nvox + 255) / 256);
    float *d_part = (float *)dh_scratch(f->dev, DH_SLOT_TMP0, sizeof(float) * 4 * nl * nblk);
    float *d_b = f->d_bnd;
    {
        dh_timed t("influence_bounds", s);
        const size_t nrb = 4 * nl * (size_t)f->g.ny * f->g.nz;
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)std::min<size_t>((nrb + 255) / 256, 4096)), dim3(256), 0, s,
                           f->d_row_bits, nrb);
        hipLaunchKernelGGL(k_layer_bounds, dim3(nblk), dim3(256), 0, s, f->g, (int)nl, f->d_layers,
                           f->d_luts, d_wet, d_part, f->d_factors, f->d_row_bits);
        DH_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_bounds, dim3((unsigned)(4 * nl)), dim3(256), 0, s, d_part, (int)nblk, d_b);
        DH_HIP(hipGetLastError());
    }
}

// Enqueues the bounds, PEAK and COUNT passes of a field and the copies of
// their results (counts, peaks, thresholds, radii) into the field's pinned
// host buffers; nothing waits.
static void count_enqueue(dh_influence *f, const float *d_wet, hipStream_t s)
{
    const int n = f->n_spots;
    if (!f->h_cnt) {
        DH_HIP(hipHostMalloc((void **)&f->h_cnt, sizeof(int) * std::max<size_t>((size_t)n * f->C, 1), 0));
        DH_HIP(hipHostMalloc((void **)&f->h_pk, sizeof(unsigned) * std::max(n, 1), 0));
        DH_HIP(hipHostMalloc((void **)&f->h_thr, sizeof(float) * std::max(n, 1), 0));
        DH_HIP(hipHostMalloc((void **)&f->h_r, sizeof(float) * std::max(n, 1), 0));
    }
    active_rows(f, d_wet, s);
    wet_bounds(f, d_wet, s);
    hipLaunchKernelGGL(k_zero_words, dim3(64), dim3(256), 0, s, f->d_peak, (size_t)std::max(n, 1));
    // Peak in two steps.  k_axis_peak evaluates the voxels nearest to each
    // spot axis; their maximum pk_a is a value of the row, so a lower bound of
    // the peak.  At lateral distance d every voxel's dose is at most
    // Pmax e^(-d^2 / 2 st_max^2) + Hmax e^(-d^2 / 2 sht_max^2) (the layer's
    // maxima), so beyond the radius where both terms are below pk_a / 2 no
    // voxel can exceed pk_a: the PEAK sweep covers that radius only, and the
    // maximum of both is the exact peak.  Radii, thresholds and the sweeps'
    // blocks are computed on the device between the passes, so the bounds,
    // peak and count passes of a field run without a host synchronisation;
    // the one synchronisation brings back the counts (the caller sizes its
    // arrays with them), the thresholds and the radii.
    if (n) {
        float hz0, hz1;
        head_depth_range(f->g, hz0, hz1);
        dh_timed t("influence_axis", s);
        hipLaunchKernelGGL(k_axis_peak, dim3((unsigned)n), dim3(64), 0, s, f->g, f->vsadx, f->vsady, n,
                           f->d_spot_layer, f->d_layers, f->d_spots, d_wet, f->d_factors, hz0, hz1, f->d_peak);
        DH_HIP(hipGetLastError());
    }
    device_radius(f, 0, s);
    device_blocks(f, s);
    launch_kernels<INF_PEAK>(f, (int)f->block_layer.size(), d_wet, nullptr, nullptr, s);
    device_radius(f, 1, s);
    device_blocks(f, s);
    launch_kernels<INF_COUNT>(f, (int)f->block_layer.size(), d_wet, nullptr, nullptr, s);
    if (n) {
        DH_HIP(hipMemcpyAsync(f->h_cnt, f->d_counts, sizeof(int) * (size_t)n * f->C, hipMemcpyDeviceToHost, s));
        DH_HIP(hipMemcpyAsync(f->h_pk, f->d_peak, sizeof(unsigned) * n, hipMemcpyDeviceToHost, s));
        DH_HIP(hipMemcpyAsync(f->h_thr, f->d_thr, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        DH_HIP(hipMemcpyAsync(f->h_r, f->d_radius, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    }
}

int64_t dh_influence_count(dh_influence *f, const float *d_wet, int64_t base, hipStream_t s);

// After count_enqueue's stream has drained: peaks, per-layer thresholds,
// the footprint check (a field whose spots reach past its footprint gets the
// full-grid WET and a recount on s, synchronised there) and the entry
// offsets (exclusive scan over (spot, chunk) in row-major order from base);
// returns nnz.
static int64_t count_finish(dh_influence *f, const float *d_wet, int64_t base, hipStream_t s)
{
    const int n = f->n_spots;
    std::vector<float> thr(f->h_thr, f->h_thr + n), r(f->h_r, f->h_r + n);
    f->peak.assign(n, 0.0f);
    for (int q = 0; q < n; q++) std::memcpy(&f->peak[q], &f->h_pk[q], 4);
    for (size_t l = 0; l < f->lthr.size(); l++) f->lthr[l] = INFINITY;
    for (int q = 0; q < n; q++)
        if (f->spot_layer[q] >= 0) f->lthr[f->spot_layer[q]] = std::min(f->lthr[f->spot_layer[q]], thr[q]);
    if (f->fp_on && d_wet == f->fp_wet && !footprint_holds(f, thr)) {
        // a spot reaches past the footprint: recompute the WET on every voxel
        f->fp_fallbacks++;
        f->fp_on = false;
        field_wet(f, f->fp_rlsp, const_cast<float *>(d_wet), nullptr, s);
        return dh_influence_count(f, d_wet, base, s);
    }
    f->offsets.assign((size_t)n * f->C + 1, 0);
    int64_t acc = base;
    for (int q = 0; q < n; q++) {
        const int l = f->spot_layer[q];
        for (int c = 0; c < f->C; c++) {
            f->offsets[(size_t)q * f->C + c] = acc;
            if (l >= 0) acc += f->h_cnt[(size_t)q * f->C + c];
        }
    }
    f->offsets[(size_t)n * f->C] = acc;
    f->nnz = acc - base;
    f->radius = r;
    f->counted_blocks = true;      // d_blocks holds the COUNT pass's blocks, d_radius its radii
    return f->nnz;
}

// PEAK + COUNT; returns nnz and fills the host offsets (exclusive scan over
// (spot, chunk) in row-major order, starting at base).
int64_t dh_influence_count(dh_influence *f, const float *d_wet, int64_t base, hipStream_t s)
{
    count_enqueue(f, d_wet, s);
    DH_HIP(hipStreamSynchronize(s));
    return count_finish(f, d_wet, base, s);
}

// Streams (default priority) of dh_influence_count_many per device, each
// with a scratch workspace of its own, and the fork event.
struct infl_pool {
    hipStream_t st[DH_INFL_STREAMS] = {};
    hipEvent_t fork = nullptr;
};

static infl_pool &infl_pool_of(int dev)
{
    static std::mutex mu;
    static infl_pool pools[64];
    DH_REQUIRE(dev >= 0 && dev < 64, "device index out of range");
    std::lock_guard<std::mutex> lk(mu);
    infl_pool &P = pools[dev];
    if (!P.fork) {
        for (hipStream_t &st : P.st) DH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        DH_HIP(hipEventCreateWithFlags(&P.fork, hipEventDisableTiming));
    }
    return P;
}

// Footprint WET + COUNT of n fields of one device, field q on stream
// q mod DH_INFL_STREAMS with that stream's workspace, so the fields' passes
// overlap on the GPU; one host synchronisation for all.  Every field's
// results equal those of dh_influence_wet + dh_influence_count on it alone;
// field q's entries start after those of fields 0..q-1 (from base).
void dh_influence_count_many(dh_influence *const *fs, int n, const float *d_rlsp, float margin_scale,
                             float *const *d_wet, int64_t base, int64_t *nnz, hipStream_t cs)
{
    if (n <= 0) return;
    const int dev = fs[0]->dev;
    infl_pool &P = infl_pool_of(dev);
    DH_HIP(hipEventRecord(P.fork, cs));
    const int used = std::min(n, DH_INFL_STREAMS);
    for (int k = 0; k < used; k++) DH_HIP(hipStreamWaitEvent(P.st[k], P.fork, 0));
    const int ws0 = dh_get_workspace();
    try {
        for (int q = 0; q < n; q++) {
            const int k = q % DH_INFL_STREAMS;
            dh_set_workspace(DH_WS_INFLUENCE + k);
            footprint_wet(fs[q], d_rlsp, margin_scale, d_wet[q], P.st[k]);
            count_enqueue(fs[q], d_wet[q], P.st[k]);
        }
    } catch (...) {
        dh_set_workspace(ws0);
        for (int k = 0; k < used; k++) (void)hipStreamSynchronize(P.st[k]);
        throw;
    }
    dh_set_workspace(ws0);
    for (int k = 0; k < used; k++) DH_HIP(hipStreamSynchronize(P.st[k]));
    int64_t b = base;
    for (int q = 0; q < n; q++) {
        nnz[q] = count_finish(fs[q], d_wet[q], b, cs);
        b += nnz[q];
    }
}

// The COUNT pass's blocks, rebuilt on the device from its radii unless
// d_blocks still holds them.
static void count_blocks(dh_influence *f, hipStream_t s)
{
    if (f->counted_blocks) return;
    const int n = f->n_spots;
    if (n) DH_HIP(hipMemcpyAsync(f->d_radius, f->radius.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
    device_blocks(f, s);
    f->counted_blocks = true;
}

void dh_influence_fill(dh_influence *f, const float *d_wet, uint32_t *d_col, float *d_val, hipStream_t s)
{
    DH_REQUIRE(f->nnz >= 0, "dh_influence_count must run before dh_influence_fill");
    const size_t n = f->offsets.size();
    if (n) DH_HIP(hipMemcpyAsync(f->d_offsets, f->offsets.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
    // the same blocks as the COUNT pass, so every entry lands in its slot;
    // the active rows and factor planes of the COUNT pass are reused: the
    // caller passes the same WET to count and fill
    count_blocks(f, s);
    launch_kernels<INF_FILL>(f, (int)f->block_layer.size(), d_wet, d_col, d_val, s);
}

void dh_influence_fill_prepare(dh_influence *f, hipStream_t s)
{
    DH_REQUIRE(f->nnz >= 0, "dh_influence_count must run before the fill is prepared");
    const size_t n = f->offsets.size();
    if (n) DH_HIP(hipMemcpyAsync(f->d_offsets, f->offsets.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s));
    count_blocks(f, s);
    DH_HIP(hipStreamSynchronize(s));
    f->fill_blocks = (int)f->block_layer.size();
}

void dh_influence_fill_enqueue(dh_influence *f, const float *d_wet, uint32_t *d_col, float *d_val, hipStream_t s)
{
    DH_REQUIRE(f->fill_blocks >= 0, "influence fill not prepared (or another pass ran since)");
    launch_kernels<INF_FILL>(f, f->fill_blocks, d_wet, d_col, d_val, s);
}

int dh_influence_device(const dh_influence *f) { return f->dev; }

int64_t dh_influence_nnz(const dh_influence *f) { return f->nnz; }

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int dh_influence_create(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                   const dh_proton_lut *lut, const float *spot_rows, int32_t n_spots,
                                   int32_t spots_per_block, int32_t row_chunks, float rel_threshold,
                                   dh_influence **out)
{
    return dh_guarded([&] {
        DH_REQUIRE(out && grid && beam && lut, "null argument");
        DH_REQUIRE(rel_threshold > 0.0f && rel_threshold < 1.0f, "threshold must be in (0, 1)");
        DH_REQUIRE(grid->nx > 0 && grid->ny > 0 && grid->nz > 0 && grid->spacing > 0.0f, "bad grid");
        DH_REQUIRE((int64_t)grid->nx * grid->ny * grid->nz < (int64_t)UINT32_MAX, "grid too large for u32 columns");
        dh_device_enter(gpu_id);
        dh_proton_tables t = dh_build_proton_tables(beam, lut, spot_rows, n_spots);
        dh_geom g = dh_make_geom(*grid, beam->iso, beam->gantry_angle, beam->couch_angle, 0.0,
                                 ((float)beam->vsadx + (float)beam->vsady) / 2.0f);
        g.sinca = 0.0f; g.cosca = 1.0f;
        dh_influence *f = dh_influence_make(gpu_id, g, (float)beam->vsadx, (float)beam->vsady, t.layers,
                                            t.luts, t.spots, n_spots, spots_per_block, row_chunks,
                                            rel_threshold);
        f->order = t.order;
        *out = f;
    });
}

extern "C" int dh_influence_wet(dh_influence *f, const float *d_rlsp, float margin_scale, float *d_wet_out,
                                void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(f && d_rlsp && d_wet_out, "null argument");
        DH_REQUIRE(margin_scale >= 0.0f, "margin_scale must be >= 0");
        dh_device_enter(f->dev);
        footprint_wet(f, d_rlsp, margin_scale, d_wet_out, (hipStream_t)stream);
    });
}

extern "C" int dh_influence_footprint(const dh_influence *f, float box[4], int32_t *active, int64_t *fallbacks)
{
    return dh_guarded([&] {
        DH_REQUIRE(f, "null handle");
        if (box) {
            box[0] = f->fp_inner.x0; box[1] = f->fp_inner.x1;
            box[2] = f->fp_inner.y0; box[3] = f->fp_inner.y1;
        }
        if (active) *active = f->fp_on ? 1 : 0;
        if (fallbacks) *fallbacks = f->fp_fallbacks;
    });
}

extern "C" int dh_influence_count(dh_influence *f, const float *d_wet, int64_t base, int64_t *nnz,
                                  int64_t *row_ptr, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(f && d_wet, "null argument");
        dh_device_enter(f->dev);
        const int64_t z = dh_influence_count(f, d_wet, base, (hipStream_t)stream);
        if (nnz) *nnz = z;
        if (row_ptr) {
            for (int q = 0; q < f->n_spots; q++) row_ptr[q] = f->offsets[(size_t)q * f->C];
            row_ptr[f->n_spots] = base + z;
        }
    });
}

extern "C" int dh_influence_count_many(dh_influence *const *fields, int32_t n, const float *d_rlsp,
                                       float margin_scale, float *const *d_wet, int64_t base, int64_t *nnz,
                                       int64_t *const *row_ptr, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(n >= 0, "field count must be >= 0");
        DH_REQUIRE(n == 0 || (fields && d_rlsp && d_wet && nnz), "null argument");
        DH_REQUIRE(margin_scale >= 0.0f, "margin_scale must be >= 0");
        if (n == 0) return;
        for (int q = 0; q < n; q++) {
            DH_REQUIRE(fields[q] && d_wet[q], "null field or WET buffer");
            DH_REQUIRE(fields[q]->dev == fields[0]->dev, "all fields must be on one device");
            for (int p = 0; p < q; p++)
                DH_REQUIRE(fields[p] != fields[q] && d_wet[p] != d_wet[q],
                           "fields and WET buffers must be distinct");
        }
        dh_device_enter(fields[0]->dev);
        dh_influence_count_many(fields, n, d_rlsp, margin_scale, d_wet, base, nnz, (hipStream_t)stream);
        if (row_ptr)
            for (int q = 0; q < n; q++) {
                const dh_influence *f = fields[q];
                if (!row_ptr[q]) continue;
                for (int r = 0; r < f->n_spots; r++) row_ptr[q][r] = f->offsets[(size_t)r * f->C];
                row_ptr[q][f->n_spots] = f->offsets[(size_t)f->n_spots * f->C];
            }
    });
}

extern "C" int dh_influence_fill(dh_influence *f, const float *d_wet, uint32_t *d_col, float *d_val,
                                 void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(f && d_wet && (f->nnz == 0 || (d_col && d_val)), "null argument");
        dh_device_enter(f->dev);
        dh_influence_fill(f, d_wet, d_col, d_val, (hipStream_t)stream);
    });
}

extern "C" int dh_influence_info(const dh_influence *f, int32_t *n_rows, int32_t *row_spot,
                                 float *row_peak)
{
    return dh_guarded([&] {
        DH_REQUIRE(f, "null handle");
        if (n_rows) *n_rows = f->n_spots;
        for (int q = 0; q < f->n_spots; q++) {
            if (row_spot) row_spot[q] = f->order[q];
            if (row_peak) row_peak[q] = q < (int)f->peak.size() ? f->peak[q] : 0.0f;
        }
    });
}

extern "C" int dh_influence_stats(const dh_influence *f, uint64_t *stats12, float *row_radius)
{
    return dh_guarded([&] {
        DH_REQUIRE(f, "null handle");
        if (stats12) {
            DH_HIP(hipSetDevice(f->dev));
            DH_HIP(hipMemcpy(stats12, f->d_stats, sizeof(uint64_t) * 12, hipMemcpyDeviceToHost));
        }
        if (row_radius)
            for (int q = 0; q < f->n_spots; q++)
                row_radius[q] = q < (int)f->radius.size() ? f->radius[q] : 0.0f;
    });
}

extern "C" int dh_influence_destroy(dh_influence *f)
{
    return dh_guarded([&] { dh_influence_free(f); });
}
