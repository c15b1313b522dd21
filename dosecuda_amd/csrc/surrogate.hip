// surrogate.hip -- BF16 MFMA kernels of the learned CCCS surrogate (a small
// 3-D convolutional encoder-decoder), plus the gamma index that scores it.
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

#include <cstring>
#include <mutex>

#include <hip/hip_bf16.h>

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 dh_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float dh_f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float dh_f32x4;

namespace {

__device__ static inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ static inline uint16_t f2bf(float f)
{   // round to nearest even: one v_cvt_pk_bf16_f32 on gfx950
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}

struct conv_args {
    const uint16_t *xa;      // source A: [N][D>>up][H>>up][W>>up][ca] bf16
    const uint16_t *xb;      // source B: [N][D][H][W][cb] bf16 or null (cb = 0)
    int ca, cb, up;          // input channel c < ca reads A (nearest-upsampled if up), else B
    const uint16_t *w;       // [taps][cout_pad][cin] bf16, cin = ca + cb
    const float *bias;       // [cout] or null
    uint16_t *y;             // [N][D][H][W][cout] bf16 or null
    float *y32;              // same, fp32, or null
    int N, D, H, W, cin, cout, cout_pad, ksize, act;
    float slope;
};

// Forward convolution, k = 1 or 3, stride 1, "same" zero padding, with the
// concatenation [upsample(A), B] of a U-Net decoder read in place.
// A workgroup of KS waves owns 32 output voxels x (32 * NCO) output channels;
// its waves split the K loop (taps x 16-channel slices) and add their FP32
// partial tiles through LDS.  Each wave keeps NCO accumulator tiles so one A
// fragment feeds NCO MFMAs, and fetches the next K step's fragments before
// issuing the current one's MFMAs.
template <int NCO, int KS>
__global__ void __launch_bounds__(64 * KS)
k_conv3d(conv_args a)
{
    extern __shared__ float red[];
    // the wave index is uniform: keep it (and the K schedule below) scalar
    const int lane = threadIdx.x & 63, wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nblk_co = a.cout_pad / (32 * NCO);
    const int cob = (int)(blockIdx.x % (unsigned)nblk_co);
    const int64_t vb = blockIdx.x / (unsigned)nblk_co;
    const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
    const int row = lane & 31, kh = lane >> 5;
    const int64_t v = vb * 32 + row;
    const bool vin = v < nvox;
    const int64_t vv = vin ? v : nvox - 1;
    const int x = (int)(vv % a.W);
    int64_t t = vv / a.W;
    const int y = (int)(t % a.H); t /= a.H;
    const int z = (int)(t % a.D);
    const int n = (int)(t / a.D);
    const int Da = a.D >> a.up, Ha = a.H >> a.up, Wa = a.W >> a.up;
    const int nsl = a.cin >> 4;
    const int kk = a.ksize * a.ksize;
    const int rad = a.ksize >> 1;
    const int KI = kk * a.ksize * nsl;
    const int k0 = (int)((int64_t)KI * wid / KS), k1 = (int)((int64_t)KI * (wid + 1) / KS);
    const int co0 = cob * 32 * NCO + row;

    dh_f32x16 acc[NCO];
#pragma unroll
    for (int j = 0; j < NCO; j++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[j][r] = 0.0f;

    // K step ki is (tap tp, 16-channel slice sl) with ki = tp * nsl + sl.
    // The fetches below run through k0, k0 + 1, ..., k1 - 1 in order, so the
    // tap offsets, slice and weight offset advance incrementally (scalar
    // state) instead of being divided out of ki at every step.
    // (element offsets fit 32 bits: make_conv bounds every tensor by 2^31)
    int f_sl, f_dx, f_dy, f_dz;
    uint32_t f_wtap;
    {
        const int tp = k0 / nsl;
        f_sl = k0 - tp * nsl;
        f_dz = tp / kk - rad;
        f_dy = (tp / a.ksize) % a.ksize - rad;
        f_dx = tp % a.ksize - rad;
        f_wtap = (uint32_t)tp * (uint32_t)(a.cout_pad * a.cin);
    }
    const uint32_t wlane = (uint32_t)co0 * (uint32_t)a.cin + (uint32_t)(kh * 8);
    const uint32_t wtap_step = (uint32_t)(a.cout_pad * a.cin);
    // fragments of the next K step: A (this lane's voxel, 8 channels) and NCO B
    auto fetch = [&](uint4 &av, uint4 *bv) {
        const int c0 = f_sl * 16;
        const int zz = z + f_dz, yy = y + f_dy, xx = x + f_dx;
        const bool ok = vin && zz >= 0 && zz < a.D && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
        av = make_uint4(0, 0, 0, 0);
        if (c0 < a.ca) {
            const uint32_t src = (((uint32_t)n * Da + (uint32_t)(zz >> a.up)) * Ha + (uint32_t)(yy >> a.up)) * Wa +
                                 (uint32_t)(xx >> a.up);
            if (ok) av = *(const uint4 *)(a.xa + (src * (uint32_t)a.ca + (uint32_t)(c0 + kh * 8)));
        } else {
            const uint32_t src = (((uint32_t)n * a.D + (uint32_t)zz) * a.H + (uint32_t)yy) * a.W + (uint32_t)xx;
            if (ok) av = *(const uint4 *)(a.xb + (src * (uint32_t)a.cb + (uint32_t)(c0 - a.ca + kh * 8)));
        }
        const uint16_t *wp = a.w + (f_wtap + wlane + (uint32_t)c0);
#pragma unroll
        for (int j = 0; j < NCO; j++) bv[j] = *(const uint4 *)(wp + (uint32_t)(j * 32 * a.cin));
        if (++f_sl == nsl) {
            f_sl = 0;
            f_wtap += wtap_step;
            if (++f_dx > rad) {
                f_dx = -rad;
                if (++f_dy > rad) {
                    f_dy = -rad;
                    ++f_dz;
                }
            }
        }
    };

    // A ring of PF K steps in flight: step ki's slot is refilled with step
    // ki + PF right after its MFMAs, so the loads of the next PF - 1 steps
    // overlap this one (the K order, hence the sum, is unchanged).
    constexpr int PF = 4;
    if (k0 < k1) {
        uint4 av[PF], bv[PF][NCO];
#pragma unroll
        for (int p = 0; p < PF; p++)
            if (k0 + p < k1) fetch(av[p], bv[p]);
        for (int kb = k0; kb < k1; kb += PF) {
#pragma unroll
            for (int p = 0; p < PF; p++) {
                const int ki = kb + p;
                if (ki < k1) {
                    dh_bf16x8 A;
                    __builtin_memcpy(&A, &av[p], 16);
#pragma unroll
                    for (int j = 0; j < NCO; j++) {
                        dh_bf16x8 B;
                        __builtin_memcpy(&B, &bv[p][j], 16);
                        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B, acc[j], 0, 0, 0);
                    }
                    if (ki + PF < k1) fetch(av[p], bv[p]);
                }
            }
        }
    }

    // D layout: lane l, register r -> row 8*(r/4) + 4*(l/32) + (r%4), column l%32
    auto emit = [&](int j, int r, int l, float s) {
        const int m = 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        const int col = cob * 32 * NCO + j * 32 + (l & 31);
        const int64_t ov = vb * 32 + m;
        if (ov >= nvox || col >= a.cout) return;
        float o = s + (a.bias ? a.bias[col] : 0.0f);
        if (a.act) o = o > 0.0f ? o : o * a.slope;
        if (a.y) a.y[ov * a.cout + col] = f2bf(o);
        if (a.y32) a.y32[ov * a.cout + col] = o;
    };
    if (KS == 1) {
#pragma unroll
        for (int j = 0; j < NCO; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) emit(j, r, lane, acc[j][r]);
        return;
    }
    constexpr int TILE = NCO * 16 * 64;
#pragma unroll
    for (int j = 0; j < NCO; j++)
#pragma unroll
        for (int r = 0; r < 16; r++) red[wid * TILE + (j * 16 + r) * 64 + lane] = acc[j][r];
    __syncthreads();
    for (int e = threadIdx.x; e < TILE; e += 64 * KS) {
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < KS; w++) s += red[w * TILE + e];
        emit(e >> 10, (e >> 6) & 15, e & 63, s);
    }
}

// 3x3x3 convolution for large volumes with few channels: a workgroup (4
// waves) owns a 2 (z) x TY (y) x 32 (x) brick of output voxels (TY = 4, or 2
// to double the grid of mid-size volumes) and 32 * NCO output channels.  Per 16-channel input slice it stages the input brick
// with its one-voxel halo (4 x 6 x 34 voxels) and the slice's weights of all
// 27 taps in LDS, then each wave runs 27 taps x 2 rows x NCO MFMAs from LDS.
// LDS layouts put the 8-channel halves of a fragment in separate planes so
// the 32 lanes of a half read 16-byte chunks at consecutive addresses.
template <int NCO, int TY>
__global__ void __launch_bounds__(256)
k_conv3d_lds(conv_args a)
{
    constexpr int TZ = 2, BY = TY + 2, BX = 34, NB = (TZ + 2) * BY * BX, MR = TZ * TY / 4;
    constexpr int NCOL = 32 * NCO;
    __shared__ __attribute__((aligned(16))) uint16_t sa[2 * NB * 8];          // [half][voxel][8]
    __shared__ __attribute__((aligned(16))) uint16_t sb[27 * 2 * NCOL * 8];   // [tap][half][col][8]
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int row = lane & 31, kh = lane >> 5;
    const int ntx = (a.W + 31) / 32, nty = (a.H + TY - 1) / TY, ntz = (a.D + TZ - 1) / TZ;
    const int ncb = a.cout_pad / NCOL;
    int64_t bi = blockIdx.x;
    const int cob = (int)(bi % ncb); bi /= ncb;
    const int tx = (int)(bi % ntx); bi /= ntx;
    const int ty = (int)(bi % nty); bi /= nty;
    const int tz = (int)(bi % ntz);
    const int n = (int)(bi / ntz);
    const int x0 = tx * 32, y0 = ty * TY, z0 = tz * TZ;
    const int Da = a.D >> a.up, Ha = a.H >> a.up, Wa = a.W >> a.up;
    const int nsl = a.cin >> 4;

    dh_f32x16 acc[MR][NCO];
#pragma unroll
    for (int m = 0; m < MR; m++)
#pragma unroll
        for (int j = 0; j < NCO; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[m][j][r] = 0.0f;

    for (int sl = 0; sl < nsl; sl++) {
        const int c0 = sl * 16;
        // brick loads first, then their LDS stores (latencies overlap)
        constexpr int NQ = (2 * NB + 255) / 256;
        uint4 stage[NQ];
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q = tid + 256 * u;
            const int vb = q >> 1, hf = q & 1;
            const int bx = vb % BX, r = vb / BX, by = r % BY, bz = r / BY;
            const int gz = z0 - 1 + bz, gy = y0 - 1 + by, gx = x0 - 1 + bx;
            stage[u] = make_uint4(0, 0, 0, 0);
            if (q < 2 * NB && gz >= 0 && gz < a.D && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                if (c0 < a.ca) {
                    const int64_t src = (((int64_t)n * Da + (gz >> a.up)) * Ha + (gy >> a.up)) * Wa + (gx >> a.up);
                    stage[u] = *(const uint4 *)(a.xa + src * a.ca + c0 + hf * 8);
                } else {
                    const int64_t src = (((int64_t)n * a.D + gz) * a.H + gy) * a.W + gx;
                    stage[u] = *(const uint4 *)(a.xb + src * a.cb + (c0 - a.ca) + hf * 8);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q = tid + 256 * u;
            if (q < 2 * NB) *(uint4 *)(sa + ((q & 1) * NB + (q >> 1)) * 8) = stage[u];
        }
        // the slice's weights, likewise
        constexpr int NW = (27 * NCOL * 2 + 255) / 256;
        uint4 wst[NW];
#pragma unroll
        for (int u = 0; u < NW; u++) {
            const int q = tid + 256 * u;
            const int r = q >> 1, col = r % NCOL, tp = r / NCOL;
            wst[u] = q < 27 * NCOL * 2
                         ? *(const uint4 *)(a.w + ((int64_t)tp * a.cout_pad + cob * NCOL + col) * a.cin + c0 + (q & 1) * 8)
                         : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < NW; u++) {
            const int q = tid + 256 * u;
            const int r = q >> 1, col = r % NCOL, tp = r / NCOL;
            if (q < 27 * NCOL * 2) *(uint4 *)(sb + ((tp * 2 + (q & 1)) * NCOL + col) * 8) = wst[u];
        }
        __syncthreads();
#pragma unroll 3
        for (int tp = 0; tp < 27; tp++) {
            const int dz = tp / 9, dy = (tp / 3) % 3, dx = tp % 3;
            dh_bf16x8 B[NCO];
#pragma unroll
            for (int j = 0; j < NCO; j++) {
                const uint4 bv = *(const uint4 *)(sb + ((tp * 2 + kh) * NCOL + j * 32 + row) * 8);
                __builtin_memcpy(&B[j], &bv, 16);
            }
#pragma unroll
            for (int m = 0; m < MR; m++) {
                const int rr = wid * MR + m, rz = rr / TY, ry = rr % TY;
                const int vb = ((rz + dz) * BY + (ry + dy)) * BX + row + dx;
                const uint4 av = *(const uint4 *)(sa + (kh * NB + vb) * 8);
                dh_bf16x8 A;
                __builtin_memcpy(&A, &av, 16);
#pragma unroll
                for (int j = 0; j < NCO; j++)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, B[j], acc[m][j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // D layout: register r of lane l -> row (x offset) 8*(r/4) + 4*(l/32) + (r%4), column l%32
#pragma unroll
    for (int m = 0; m < MR; m++) {
        const int rr = wid * MR + m;
        const int gz = z0 + rr / TY, gy = y0 + rr % TY;
        if (gz >= a.D || gy >= a.H) continue;
#pragma unroll
        for (int j = 0; j < NCO; j++) {
            const int col = cob * NCOL + j * 32 + row;
            if (col >= a.cout) continue;
            const float bsv = a.bias ? a.bias[col] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int gx = x0 + 8 * (r >> 2) + 4 * kh + (r & 3);
                if (gx >= a.W) continue;
                const int64_t ov = (((int64_t)n * a.D + gz) * a.H + gy) * a.W + gx;
                float o = acc[m][j][r] + bsv;
                if (a.act) o = o > 0.0f ? o : o * a.slope;
                if (a.y) a.y[ov * a.cout + col] = f2bf(o);
                if (a.y32) a.y32[ov * a.cout + col] = o;
            }
        }
    }
}

// 3x3x3 convolution with at most 16 output channels (the full-resolution
// layers), on v_mfma_f32_16x16x32_bf16.  Operand layout: lane l holds A row
// (l mod 16) and B column (l mod 16), each with the 8 consecutive K elements
// 8*(l/16) .. 8*(l/16)+7; the result lane l holds column (l mod 16), rows
// 4*(l/16) .. 4*(l/16)+3.  A K step is two taps x 16 channels (K element k:
// tap 2p + k/16, channel k mod 16), so a 16-channel slice takes 14 MFMAs per
// tile of 16 voxels (the 14th pairs tap 26 with a zero weight), with no
// padded output columns.  A workgroup (4 waves) owns a 2 (z) x TY (y) x 32 (x)
// brick; per slice it stages the input brick with its one-voxel halo in LDS
// (two 8-channel planes, so 16 lanes read consecutive 16-byte chunks), each
// lane loads its 14 weight fragments into registers, and each wave runs its
// TY/2 rows x 2 tiles x 14 MFMAs.
template <int TY>
__global__ void __launch_bounds__(256)
k_conv3d_c16(conv_args a)
{
    constexpr int TZ = 2, BY = TY + 2, BX = 34, NB = (TZ + 2) * BY * BX;
    // the two 8-channel planes, the second 128 bytes further on, so that the
    // 16 lanes of a staging store (8 per plane) fall on different banks
    constexpr int PL = NB * 8 + 64;
    constexpr int RPW = TZ * TY / 4;                 // rows of 32 voxels per wave
    __shared__ __attribute__((aligned(16))) uint16_t sa[2 * PL];   // [half][voxel][8]
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const int half = kq & 1, tsub = kq >> 1;
    const int ntx = (a.W + 31) / 32, nty = (a.H + TY - 1) / TY, ntz = (a.D + TZ - 1) / TZ;
    uint32_t bi = blockIdx.x;
    const int tx = (int)(bi % (uint32_t)ntx); bi /= (uint32_t)ntx;
    const int ty = (int)(bi % (uint32_t)nty); bi /= (uint32_t)nty;
    const int tz = (int)(bi % (uint32_t)ntz);
    const int n = (int)(bi / (uint32_t)ntz);
    const int x0 = tx * 32, y0 = ty * TY, z0 = tz * TZ;
    const int Da = a.D >> a.up, Ha = a.H >> a.up, Wa = a.W >> a.up;
    const int nsl = a.cin >> 4;
    // element indices in 32 bits: the host admits tensors of fewer than 2^31
    // elements

    dh_f32x4 acc[RPW][2];
#pragma unroll
    for (int m = 0; m < RPW; m++)
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++)
#pragma unroll
            for (int r = 0; r < 4; r++) acc[m][h2][r] = 0.0f;

    for (int sl = 0; sl < nsl; sl++) {
        const int c0 = sl * 16;
        // all of a thread's brick loads are issued before any LDS store, so
        // their latencies overlap instead of adding up
        constexpr int NQ = (2 * NB + 255) / 256;
        uint4 stage[NQ];
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q = tid + 256 * u;
            const int vb = q >> 1, hf = q & 1;
            const int bx = vb % BX, r = vb / BX, by = r % BY, bz = r / BY;
            const int gz = z0 - 1 + bz, gy = y0 - 1 + by, gx = x0 - 1 + bx;
            stage[u] = make_uint4(0, 0, 0, 0);
            if (q < 2 * NB && gz >= 0 && gz < a.D && gy >= 0 && gy < a.H && gx >= 0 && gx < a.W) {
                if (c0 < a.ca) {
                    const uint32_t src = ((uint32_t)(n * Da + (gz >> a.up)) * (uint32_t)Ha + (uint32_t)(gy >> a.up)) *
                                             (uint32_t)Wa + (uint32_t)(gx >> a.up);
                    stage[u] = *(const uint4 *)(a.xa + (src * (uint32_t)a.ca + (uint32_t)(c0 + hf * 8)));
                } else {
                    const uint32_t src = ((uint32_t)(n * a.D + gz) * (uint32_t)a.H + (uint32_t)gy) * (uint32_t)a.W +
                                         (uint32_t)gx;
                    stage[u] = *(const uint4 *)(a.xb + (src * (uint32_t)a.cb + (uint32_t)(c0 - a.ca + hf * 8)));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < NQ; u++) {
            const int q = tid + 256 * u;
            if (q < 2 * NB) *(uint4 *)(sa + (q & 1) * PL + (q >> 1) * 8) = stage[u];
        }
        dh_bf16x8 B[14];
        const uint16_t *wl = a.w + ((uint32_t)col * (uint32_t)a.cin + (uint32_t)(c0 + half * 8));
        const uint32_t wtap = (uint32_t)a.cout_pad * (uint32_t)a.cin;
#pragma unroll
        for (int p = 0; p < 14; p++) {
            const int tp = 2 * p + tsub;
            uint4 bv = make_uint4(0, 0, 0, 0);
            if (tp < 27 && col < a.cout) bv = *(const uint4 *)(wl + (uint32_t)tp * wtap);
            __builtin_memcpy(&B[p], &bv, 16);
        }
        __syncthreads();
        const uint16_t *sh = sa + half * PL + col * 8;
#pragma unroll
        for (int p = 0; p < 14; p++) {
            const int tp = min(2 * p + tsub, 26);
            const int dz = tp / 9, dy = (tp / 3) % 3, dx = tp % 3;
#pragma unroll
            for (int m = 0; m < RPW; m++) {
                const int rr = wid * RPW + m, rz = rr / TY, ry = rr % TY;
#pragma unroll
                for (int h2 = 0; h2 < 2; h2++) {
                    const int vb = ((rz + dz) * BY + (ry + dy)) * BX + h2 * 16 + dx;
                    const uint4 av = *(const uint4 *)(sh + vb * 8);
                    dh_bf16x8 A;
                    __builtin_memcpy(&A, &av, 16);
                    acc[m][h2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B[p], acc[m][h2], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    if (col >= a.cout) return;
    const float bsv = a.bias ? a.bias[col] : 0.0f;
    const int nx = min(32, a.W - x0);                // valid voxels of the brick's rows
    // the bias and the activation of every output first (branch-free), then
    // the stores: the common case (whole rows of 32 voxels, 16 channels, bf16
    // output only: the network's hidden layers) without per-element tests and
    // with constant store offsets
#pragma unroll
    for (int m = 0; m < RPW; m++)
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float o = acc[m][h2][r] + bsv;
                acc[m][h2][r] = a.act ? (o > 0.0f ? o : o * a.slope) : o;
            }
    const bool fast = nx == 32 && a.cout == 16 && a.y && !a.y32;
#pragma unroll
    for (int m = 0; m < RPW; m++) {
        const int rr = wid * RPW + m;
        const int gz = z0 + rr / TY, gy = y0 + rr % TY;
        if (gz >= a.D || gy >= a.H) continue;
        // element of (voxel x0 + 4 kq of this row, channel col)
        const uint32_t e0 = (((uint32_t)(n * a.D + gz) * (uint32_t)a.H + (uint32_t)gy) * (uint32_t)a.W +
                             (uint32_t)(x0 + 4 * kq)) * (uint32_t)a.cout + (uint32_t)col;
        if (fast) {
            uint16_t *yr = a.y + e0;
#pragma unroll
            for (int h2 = 0; h2 < 2; h2++)
#pragma unroll
                for (int r = 0; r < 4; r++) yr[(16 * h2 + r) * 16] = f2bf(acc[m][h2][r]);
            continue;
        }
#pragma unroll
        for (int h2 = 0; h2 < 2; h2++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (16 * h2 + 4 * kq + r >= nx) continue;
                const uint32_t off = (uint32_t)(16 * h2 + r) * (uint32_t)a.cout;
                if (a.y) a.y[e0 + off] = f2bf(acc[m][h2][r]);
                if (a.y32) a.y32[e0 + off] = acc[m][h2][r];
            }
    }
}

// 1x1x1 convolution with at most 16 output channels and one source (the
// output layer), on v_mfma_f32_16x16x32_bf16 with the operand layout above:
// a wave owns 64 consecutive voxels (4 tiles of 16), the K loop runs over
// 32-channel steps (channels past cin are zero), every fragment is one
// 16-byte load.
__global__ void __launch_bounds__(256)
k_conv1_c16(conv_args a)
{
    const int lane = threadIdx.x & 63, col = lane & 15, kq = lane >> 4;
    const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
    const int64_t v0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
    if (v0 >= nvox) return;
    dh_f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = 0.0f;
    for (int c0 = 0; c0 < a.cin; c0 += 32) {
        const int ci = c0 + 8 * kq;
        uint4 bv = make_uint4(0, 0, 0, 0);
        if (ci < a.cin && col < a.cout) bv = *(const uint4 *)(a.w + (int64_t)col * a.cin + ci);
        dh_bf16x8 B;
        __builtin_memcpy(&B, &bv, 16);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int64_t v = v0 + 16 * t + col;
            uint4 av = make_uint4(0, 0, 0, 0);
            if (ci < a.cin && v < nvox) av = *(const uint4 *)(a.xa + v * a.cin + ci);
            dh_bf16x8 A;
            __builtin_memcpy(&A, &av, 16);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
        }
    }
    if (col >= a.cout) return;
    const float bsv = a.bias ? a.bias[col] : 0.0f;
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int64_t v = v0 + 16 * t + 4 * kq + r;
            if (v >= nvox) continue;
            float o = acc[t][r] + bsv;
            if (a.act) o = o > 0.0f ? o : o * a.slope;
            if (a.y) a.y[v * a.cout + col] = f2bf(o);
            if (a.y32) a.y32[v * a.cout + col] = o;
        }
}

// avg-pool 2x2x2 on channels-last bf16 [N][D][H][W][C] -> [N][D/2][H/2][W/2][C]
__global__ void k_pool2(const uint16_t *__restrict__ x, int N, int D, int H, int W, int C,
                        uint16_t *__restrict__ y)
{
    const int d = D >> 1, h = H >> 1, w = W >> 1;
    const int64_t n = (int64_t)N * d * h * w * C;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        int64_t r = i / C;
        const int xx = (int)(r % w); r /= w;
        const int yy = (int)(r % h); r /= h;
        const int zz = (int)(r % d);
        const int nn = (int)(r / d);
        float s = 0.0f;
#pragma unroll
        for (int o = 0; o < 8; o++) {
            const int64_t src = ((((int64_t)nn * D + 2 * zz + (o >> 2)) * H + 2 * yy + ((o >> 1) & 1)) * W +
                                 2 * xx + (o & 1)) * C + c;
            s += bf2f(x[src]);
        }
        y[i] = f2bf(s * 0.125f);
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

// ---- surrogate input / output and the checked fallback ----------------------

// One atomic max per workgroup (256 threads) of non-negative floats, via
// their ordered bit patterns.
__device__ static inline void block_max_to(float m, float *out)
{
    __shared__ float s_m[4];
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if (__lane_id() == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]));
        // the maximum only grows, so a value read at any time bounds it from
        // below: a workgroup whose maximum does not exceed it has nothing to add
        if (b > 0.0f && b > *(volatile float *)out) atomicMax((unsigned int *)out, __float_as_uint(b));
    }
}

// max of a non-negative volume (16-byte loads when the base allows)
__device__ static inline void max_nonneg(const float *__restrict__ v, int64_t n, float *__restrict__ out)
{
    float m = 0.0f;
    const int64_t n4 = ((uintptr_t)v & 15) == 0 ? n / 4 : 0;
    const float4 *v4 = (const float4 *)v;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4; i += step) {
        const float4 q = v4[i];
        m = fmaxf(m, fmaxf(fmaxf(q.x, q.y), fmaxf(q.z, q.w)));
    }
    for (int64_t i = 4 * n4 + blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += step) m = fmaxf(m, v[i]);
    block_max_to(m, out);
}

__global__ void __launch_bounds__(256)
k_max_nonneg(const float *__restrict__ v, int64_t n, float *__restrict__ out)
{
    max_nonneg(v, n, out);
}

// Per-point pointers of a batch of control points (blockIdx.y = point).
constexpr int kPackBatchMax = 8;
struct sur_pack_batch {
    const float *terma[kPackBatchMax];
    float *tmax[kPackBatchMax];
    uint16_t *x[kPackBatchMax];
    float dir[kPackBatchMax][3];
};

__global__ void __launch_bounds__(256)
k_max_nonneg_batch(sur_pack_batch b, int64_t n)
{
    max_nonneg(b.terma[blockIdx.y], n, b.tmax[blockIdx.y]);
}

__global__ void k_fill_f32(float *p, int n, float v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_fill_i32(int *p, int n, int v)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// Input cells on the half-resolution grid of the 16-padded volume: channels
// 0-7 TERMA / max, 8-15 density of the 2x2x2 sub-voxels (sub = 4 dz + 2 dy +
// dx), 16-18 the beam direction, 19 the spacing channel, 20-31 zero.  One
// thread writes one cell's 32 channels (four 16-byte stores).
__device__ static inline void sur_pack(const float *__restrict__ terma, const float *__restrict__ dens, int nz,
                                       int ny, int nx, int cz, int cy, int cx, const float *__restrict__ tmax,
                                       float d0, float d1, float d2, float spc, uint16_t *__restrict__ out)
{
    const int64_t n = (int64_t)cz * cy * cx;
    const float scale = fmaxf(*tmax, 1e-12f);
    uint4 cst;
    {
        uint16_t h[8] = {f2bf(d0), f2bf(d1), f2bf(d2), f2bf(spc), 0, 0, 0, 0};
        __builtin_memcpy(&cst, h, 16);
    }
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(c % cx);
        const int64_t r = c / cx;
        const int cj = (int)(r % cy), ck = (int)(r / cy);
        uint16_t ht[8], hd[8];
#pragma unroll
        for (int sub = 0; sub < 8; sub++) {
            const int k = 2 * ck + (sub >> 2), j = 2 * cj + ((sub >> 1) & 1), ii = 2 * ci + (sub & 1);
            const bool in = k < nz && j < ny && ii < nx;
            const int64_t src = ((int64_t)k * ny + j) * nx + ii;
            ht[sub] = f2bf(in ? terma[src] / scale : 0.0f);
            hd[sub] = f2bf(in ? dens[src] : DH_AIR_DENSITY);
        }
        uint4 a, b;
        __builtin_memcpy(&a, ht, 16);
        __builtin_memcpy(&b, hd, 16);
        uint4 *o = (uint4 *)(out + c * 32);
        o[0] = a;
        o[1] = b;
        o[2] = cst;
        o[3] = make_uint4(0, 0, 0, 0);
    }
}

__global__ void __launch_bounds__(256)
k_sur_pack(const float *__restrict__ terma, const float *__restrict__ dens, int nz, int ny,
           int nx, int cz, int cy, int cx, const float *__restrict__ tmax, float d0,
           float d1, float d2, float spc, uint16_t *__restrict__ out)
{
    sur_pack(terma, dens, nz, ny, nx, cz, cy, cx, tmax, d0, d1, d2, spc, out);
}

__global__ void __launch_bounds__(256)
k_sur_pack_batch(sur_pack_batch b, const float *__restrict__ dens, int nz, int ny, int nx, int cz, int cy, int cx,
                 float spc)
{
    const int q = blockIdx.y;
    sur_pack(b.terma[q], dens, nz, ny, nx, cz, cy, cx, b.tmax[q], b.dir[q][0], b.dir[q][1], b.dir[q][2], spc,
             b.x[q]);
}

// Surrogate dose: the network's sub-voxel outputs (fp32 [cells][16],
// channels 0-7) back on the voxel grid, scaled by max TERMA and the
// network's output scale, plus the electron term, times MU calibration;
// zero where TERMA <= 0.01 or the value is not positive.  Also the dose max.
// One thread per cell of the volume (two 16-byte loads of its outputs).
__device__ static inline void sur_finish(const float *__restrict__ y32, int cy, int cx,
                                         const float *__restrict__ tmax, float out_scale,
                                         const float *__restrict__ terma, const float *__restrict__ electron,
                                         float mu_scale, int nz, int ny, int nx, float *__restrict__ dose,
                                         float *__restrict__ dmax)
{
    const int vz = (nz + 1) / 2, vy = (ny + 1) / 2, vx = (nx + 1) / 2;
    const int64_t n = (int64_t)vz * vy * vx;
    const float s = fmaxf(*tmax, 1e-12f) * out_scale;
    float m = 0.0f;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int ci = (int)(c % vx);
        const int64_t r = c / vx;
        const int cj = (int)(r % vy), ck = (int)(r / vy);
        const int64_t cc = ((int64_t)ck * cy + cj) * cx + ci;
        const float4 y0 = *(const float4 *)(y32 + cc * 16), y1 = *(const float4 *)(y32 + cc * 16 + 4);
        const float yy[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
#pragma unroll
        for (int sub = 0; sub < 8; sub++) {
            const int k = 2 * ck + (sub >> 2), j = 2 * cj + ((sub >> 1) & 1), i = 2 * ci + (sub & 1);
            if (k >= nz || j >= ny || i >= nx) continue;
            const int64_t v = ((int64_t)k * ny + j) * nx + i;
            const float sup = yy[sub] * s;
            const float d = (sup + electron[v]) * mu_scale;
            const float o = (terma[v] > 0.01f && d > 0.0f) ? d : 0.0f;
            dose[v] = o;
            m = fmaxf(m, o);
        }
    }
    block_max_to(m, dmax);
}

__global__ void __launch_bounds__(256)
k_sur_finish(const float *__restrict__ y32, int cy, int cx, const float *__restrict__ tmax,
             float out_scale, const float *__restrict__ terma, const float *__restrict__ electron,
             float mu_scale, int nz, int ny, int nx, float *__restrict__ dose,
             float *__restrict__ dmax)
{
    sur_finish(y32, cy, cx, tmax, out_scale, terma, electron, mu_scale, nz, ny, nx, dose, dmax);
}

// Per-point buffers of the batched check (blockIdx.y = point).  work: int32
// {lattice count, fallback count, 2 spare}, then per lattice cell lat_ok,
// bad and redo, then the lattice list (one per cell) and the fallback list
// (one per voxel).
struct sur_check_batch {
    const float *y32[kPackBatchMax], *tmax[kPackBatchMax];
    float mu_scale[kPackBatchMax];
    const float *terma[kPackBatchMax], *electron[kPackBatchMax];
    float *dose[kPackBatchMax], *dmax[kPackBatchMax], *plain[kPackBatchMax];
    int *work[kPackBatchMax];
};

static_assert(sizeof(sur_check_batch) <= 4096, "check batch arguments exceed the kernel argument space");

__device__ static inline int *chk_lat_ok(int *w) { return w + 4; }
__device__ static inline int *chk_bad(int *w, int64_t nc) { return w + 4 + nc; }
__device__ static inline int *chk_redo(int *w, int64_t nc) { return w + 4 + 2 * nc; }
__device__ static inline int *chk_lat_list(int *w, int64_t nc) { return w + 4 + 3 * nc; }
__device__ static inline int *chk_fb_list(int *w, int64_t nc) { return w + 4 + 4 * nc; }

// The surrogate dose of each point of a batch (network outputs y32 at
// cells * 16 floats per point), as k_sur_finish: workgroups take rows of
// cells (ck, cj) of their point (blockIdx.y) in turn, staging 64 cells'
// eight sub-voxel outputs in LDS at a time so that the four voxel rows are
// read and written with consecutive lanes on consecutive voxels.  Block (0, 0) of each point
// also zeroes the point's check counters.
__global__ void __launch_bounds__(256)
k_sur_finish_batch(sur_check_batch b, int cy, int cx, float out_scale, int nz, int ny, int nx)
{
    __shared__ float s_y[64][9];
    const int q = blockIdx.y;
    const int vy = (ny + 1) / 2, nrow = ((nz + 1) / 2) * vy;
    if (blockIdx.x == 0 && threadIdx.x < 4) b.work[q][threadIdx.x] = 0;
    const float *__restrict__ y32 = b.y32[q];
    const float *__restrict__ terma = b.terma[q];
    const float *__restrict__ electron = b.electron[q];
    float *__restrict__ dose = b.dose[q];
    const float s = fmaxf(*b.tmax[q], 1e-12f) * out_scale;
    const float mu_scale = b.mu_scale[q];
    const int t = threadIdx.x, row = t >> 6, lane = t & 63;
    const int dz = row >> 1, dy = row & 1;
    float m = 0.0f;
    for (int cr = blockIdx.x; cr < nrow; cr += gridDim.x) {
    const int ck = cr / vy, cj = cr % vy;
    const int k = 2 * ck + dz, j = 2 * cj + dy;
    const bool in_row = k < nz && j < ny;
    const int64_t vrow = ((int64_t)k * ny + j) * nx;
    const int64_t crow = ((int64_t)ck * cy + cj) * cx;
    for (int c0 = 0; 2 * c0 < nx; c0 += 64) {
        {
            // 64 cells x 8 outputs: four lanes per cell, two floats each
            const int cell = t >> 2, part = t & 3;
            float2 y = make_float2(0.0f, 0.0f);
            if (c0 + cell < cx) y = *(const float2 *)(y32 + (crow + c0 + cell) * 16 + 2 * part);
            s_y[cell][2 * part] = y.x;
            s_y[cell][2 * part + 1] = y.y;
        }
        __syncthreads();
        if (in_row) {
            for (int xx = lane; xx < 128; xx += 64) {
                const int x = 2 * c0 + xx;
                if (x >= nx) break;
                const int64_t v = vrow + x;
                const float sup = s_y[xx >> 1][4 * dz + 2 * dy + (x & 1)] * s;
                const float d = (sup + electron[v]) * mu_scale;
                const float o = (terma[v] > 0.01f && d > 0.0f) ? d : 0.0f;
                dose[v] = o;
                m = fmaxf(m, o);
            }
        }
        __syncthreads();
    }
    }
    block_max_to(m, b.dmax[q]);
}

// The network's output layer (k_conv1_c16: 1x1x1, at most 16 outputs, one
// source, channels 0-7 the sub-voxel outputs) fused with the surrogate dose
// of a batch of points (the network's cells of point q are cells q * cells
// .. (q + 1) * cells - 1): each wave forms its 64 cells' outputs with the
// layer's MFMAs, the workgroup's 256 cells' eight sub-voxel outputs go
// through LDS, and the workgroup then forms the doses of those cells' voxels
// as k_sur_finish does (consecutive lanes on consecutive voxels of a row) and
// adds its maximum to the point's.  The values are exactly those of the
// layer's fp32 output fed to k_sur_finish.  The first workgroup of each point
// zeroes the point's check counters.
__global__ void __launch_bounds__(256)
k_conv1_finish(conv_args a, sur_check_batch b, float out_scale, int nz, int ny, int nx, int cy, int cx,
               int64_t cells)
{
    __shared__ float s_y[256][9];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, col = lane & 15, kq = lane >> 4;
    const int64_t cb0 = (int64_t)blockIdx.x * 256;          // first cell of the workgroup (all points)
    const int q = (int)(cb0 / cells);
    const int64_t v0 = cb0 + wid * 64;
    dh_f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) acc[t][r] = 0.0f;
    for (int c0 = 0; c0 < a.cin; c0 += 32) {
        const int ci = c0 + 8 * kq;
        uint4 bv = make_uint4(0, 0, 0, 0);
        if (ci < a.cin && col < a.cout) bv = *(const uint4 *)(a.w + (int64_t)col * a.cin + ci);
        dh_bf16x8 B;
        __builtin_memcpy(&B, &bv, 16);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int64_t v = v0 + 16 * t + col;
            uint4 av = make_uint4(0, 0, 0, 0);
            if (ci < a.cin) av = *(const uint4 *)(a.xa + v * a.cin + ci);
            dh_bf16x8 A;
            __builtin_memcpy(&A, &av, 16);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[t], 0, 0, 0);
        }
    }
    if (col < 8) {                       // the eight sub-voxel outputs
        const float bsv = a.bias ? a.bias[col] : 0.0f;
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float o = acc[t][r] + bsv;
                if (a.act) o = o > 0.0f ? o : o * a.slope;
                s_y[wid * 64 + 16 * t + 4 * kq + r][col] = o;
            }
    }
    if (cb0 % cells == 0 && tid < 4) b.work[q][tid] = 0;
    __syncthreads();
    const float *__restrict__ terma = b.terma[q];
    const float *__restrict__ electron = b.electron[q];
    float *__restrict__ dose = b.dose[q];
    const float s = fmaxf(*b.tmax[q], 1e-12f) * out_scale;
    const float mu_scale = b.mu_scale[q];
    // cell and voxel indices in 32 bits (a point's volume and cells number
    // fewer than 2^31)
    const uint32_t cl0 = (uint32_t)(cb0 - (int64_t)q * cells);    // first cell within the point
    float m = 0.0f;
    for (int idx = tid; idx < 2048; idx += 256) {
        // voxel rows (dz, dy) outermost, then cells, then dx: consecutive
        // lanes on consecutive voxels of a row
        const int row = idx >> 9, cell = (idx & 511) >> 1, dx = idx & 1;
        const int dz = row >> 1, dy = row & 1;
        const uint32_t c = cl0 + (uint32_t)cell, cr = c / (uint32_t)cx;
        const int ci = (int)(c - cr * (uint32_t)cx), cj = (int)(cr % (uint32_t)cy), ck = (int)(cr / (uint32_t)cy);
        const int k = 2 * ck + dz, j = 2 * cj + dy, i = 2 * ci + dx;
        if (k >= nz || j >= ny || i >= nx) continue;
        const uint32_t v = ((uint32_t)k * (uint32_t)ny + (uint32_t)j) * (uint32_t)nx + (uint32_t)i;
        const float sup = s_y[cell][4 * dz + 2 * dy + dx] * s;
        const float d = (sup + electron[v]) * mu_scale;
        const float o = (terma[v] > 0.01f && d > 0.0f) ? d : 0.0f;
        dose[v] = o;
        m = fmaxf(m, o);
    }
    block_max_to(m, b.dmax[q]);
}

// block-aggregated append of the flagged items of this block to list
__device__ static inline void block_append(bool flag, int item, int *__restrict__ list, int *__restrict__ counter)
{
    __shared__ int s_cnt[4];
    __shared__ int s_base;
    const uint64_t mk = __builtin_amdgcn_ballot_w64(flag);
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    if (lane == 0) s_cnt[wave] = __popcll(mk);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int w = 0; w < 4; w++) { const int c = s_cnt[w]; s_cnt[w] = tot; tot += c; }
        s_base = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    if (flag) list[s_base + s_cnt[wave] + __popcll(mk & ((1ull << lane) - 1ull))] = item;
}

// The check voxel of each lattice cell (stride^3 voxels from its low corner):
// the low corner when its TERMA > 0.01, else the cell's centre voxel (offset
// stride / 2 on each axis) when that lies in the volume with TERMA > 0.01, so
// cells across the field edge are checked rather than all recomputed.
// lat_ok: 1 corner, 2 centre, 0 no check voxel.
__device__ static inline int64_t sur_check_voxel(int code, int ck, int cj, int ci, int stride, int ny, int nx)
{
    const int h = code == 2 ? stride / 2 : 0;
    return ((int64_t)(ck * stride + h) * ny + (cj * stride + h)) * nx + (ci * stride + h);
}

__global__ void __launch_bounds__(256)
k_sur_lattice(sur_check_batch bt, int nz, int ny, int nx, int stride, int lz, int ly, int lx)
{
    const int q = blockIdx.y;
    const float *__restrict__ terma = bt.terma[q];
    int *w = bt.work[q];
    const int64_t nc = (int64_t)lz * ly * lx;
    const int64_t c = blockIdx.x * (int64_t)256 + threadIdx.x;
    bool act = false;
    int v = 0;
    if (c < nc) {
        const int ci = (int)(c % lx);
        const int64_t r = c / lx;
        const int cj = (int)(r % ly), ck = (int)(r / ly);
        const int h = stride / 2;
        int code = 0;
        if (!(terma[sur_check_voxel(1, ck, cj, ci, stride, ny, nx)] <= 0.01f)) code = 1;
        else if (h > 0 && ck * stride + h < nz && cj * stride + h < ny && ci * stride + h < nx &&
                 !(terma[sur_check_voxel(2, ck, cj, ci, stride, ny, nx)] <= 0.01f)) code = 2;
        act = code != 0;
        v = (int)sur_check_voxel(code ? code : 1, ck, cj, ci, stride, ny, nx);
        chk_lat_ok(w)[c] = code;
    }
    block_append(act, v, chk_lat_list(w, nc), w);
}

// per lattice cell: the surrogate missed the plain value at the cell's check
// voxel by more than tol * max
__global__ void __launch_bounds__(256)
k_sur_flag(sur_check_batch bt, float tol, int ny, int nx, int stride, int lz, int ly, int lx)
{
    const int q = blockIdx.y;
    int *w = bt.work[q];
    const int64_t nc = (int64_t)lz * ly * lx;
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int ci = (int)(c % lx);
    const int64_t r = c / lx;
    const int cj = (int)(r % ly), ck = (int)(r / ly);
    const int code = chk_lat_ok(w)[c];
    const float thr = tol * fmaxf(*bt.dmax[q], 1e-12f);
    const int64_t v = sur_check_voxel(code ? code : 1, ck, cj, ci, stride, ny, nx);
    chk_bad(w, nc)[c] = (code && fabsf(bt.dose[q][v] - bt.plain[q][v]) > thr) ? 1 : 0;
}

// per cell: recompute its voxels when it has no check voxel (neither its
// corner nor its centre has TERMA > 0.01) or the cell failed, or (dilate 1) one of
// its 6 face neighbours or (dilate 2) one of its 26 neighbours failed
__global__ void __launch_bounds__(256)
k_sur_dilate(sur_check_batch bt, int lz, int ly, int lx, int dilate)
{
    const int q = blockIdx.y;
    int *w = bt.work[q];
    const int64_t nc = (int64_t)lz * ly * lx;
    const int *__restrict__ lat_ok = chk_lat_ok(w);
    const int *__restrict__ bad = chk_bad(w, nc);
    const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int ci = (int)(c % lx);
    const int64_t r = c / lx;
    const int cj = (int)(r % ly), ck = (int)(r / ly);
    bool fb = !lat_ok[c] || bad[c] != 0;
    for (int dk = -1; dk <= 1 && !fb && dilate > 0; dk++)
        for (int dj = -1; dj <= 1 && !fb; dj++)
            for (int di = -1; di <= 1 && !fb; di++) {
                if (dilate == 1 && (dk != 0) + (dj != 0) + (di != 0) != 1) continue;
                const int a = ck + dk, b = cj + dj, e = ci + di;
                if (a < 0 || a >= lz || b < 0 || b >= ly || e < 0 || e >= lx) continue;
                fb = bad[((int64_t)a * ly + b) * lx + e] != 0;
            }
    chk_redo(w, nc)[c] = fb ? 1 : 0;
}

// voxels with TERMA > 0.01 in cells marked for recomputation, listed brick
// by brick: a workgroup walks one 16 x 16 x 8 brick (x fastest), so the 64
// consecutive list entries of a cone work item lie close together in 3-D
// (in voxel order they would span whole rows, often both edges of a field)
constexpr int kFallbackRounds = 8;
constexpr int kFbBx = 16, kFbBy = 16, kFbBz = 8;
static_assert(kFbBx * kFbBy * kFbBz == 256 * kFallbackRounds, "one brick per workgroup");

__global__ void __launch_bounds__(256)
k_sur_fallback(sur_check_batch bt, int band, int nz, int ny, int nx, int stride, int lz, int ly, int lx)
{
    const int q = blockIdx.y;
    int *w = bt.work[q];
    const int64_t nc = (int64_t)lz * ly * lx;
    const float *__restrict__ terma = bt.terma[q];
    const float *__restrict__ electron = bt.electron[q];
    const int *__restrict__ redo = chk_redo(w, nc);
    const int nbx = (nx + kFbBx - 1) / kFbBx, nby = (ny + kFbBy - 1) / kFbBy;
    // brick blockIdx.x: its voxels are items blockIdx.x * 2048 + (0 .. 2047)
    const int bx = (int)(blockIdx.x % (uint32_t)nbx), rb = (int)(blockIdx.x / (uint32_t)nbx);
    const int by = rb % nby, bz = rb / nby;
    const uint32_t base = blockIdx.x * (uint32_t)(256 * kFallbackRounds);
    auto voxel = [&](uint32_t u) -> int {
        const int w = (int)(u - base);
        const int i = bx * kFbBx + w % kFbBx, j = by * kFbBy + (w / kFbBx) % kFbBy, k = bz * kFbBz + w / (kFbBx * kFbBy);
        return i < nx && j < ny && k < nz ? i + nx * (j + ny * k) : -1;
    };
    dh_block_compact_map<kFallbackRounds>(base, 0xffffffffu, voxel, [&](uint32_t v) {
        if (terma[v] <= 0.01f) return false;
        // the build-up band: a nonzero electron-contamination term means
        // WET < electron_fitted_dmax (IMRTClasses.cu:238-243)
        if (band && electron[v] > 0.0f) return true;
        const int i = (int)(v % (uint32_t)nx);
        const uint32_t r = v / (uint32_t)nx;
        const int j = (int)(r % (uint32_t)ny), k = (int)(r / (uint32_t)ny);
        return redo[((int64_t)(k / stride) * ly + j / stride) * lx + i / stride] != 0;
    }, chk_fb_list(w, nc), w + 1);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

template <int NCO, int KS>
void launch_conv(const conv_args &a, int64_t items, hipStream_t s)
{
    const size_t lds = KS > 1 ? (size_t)KS * NCO * 16 * 64 * sizeof(float) : 0;
    hipLaunchKernelGGL((k_conv3d<NCO, KS>), dim3((unsigned)items), dim3(64 * KS), lds, s, a);
}

// (NCO, KS) for a layer: the widest output tile that still gives a few
// thousand waves, then split K until the grid fills the chip (at least
// eight K steps per wave; LDS of the partial tiles <= 64 KiB).
void conv_shape(int64_t nvox, int cout_pad, int ki, int &nco, int &ks)
{
    const int64_t nvb = (nvox + 31) / 32;
    const int ntile = cout_pad / 32;
    nco = 1;
    for (int c : {4, 2}) {
        if (ntile % c == 0 && nvb * (ntile / c) >= 2048) { nco = c; break; }
    }
    ks = 1;
    while (ks < 16 && nvb * (ntile / nco) * ks < 8192 && ki / (2 * ks) >= 8 && 2 * ks * nco <= 16) ks *= 2;
}

}  // namespace

namespace {

void launch_pool(const uint16_t *x, int n, int d, int h, int w, int c, uint16_t *y, hipStream_t s);

// Picks the kernel variant (explicit nco / ksplit, or the library's choice)
// and launches it; throws dh_error on bad arguments.
std::string conv_key(const conv_args &a)
{
    char b[160];
    snprintf(b, sizeof(b), "%d,%d,%d,%d,%d,%d,%d,%d,%d", a.N, a.D, a.H, a.W, a.ca, a.cb, a.up, a.cout, a.ksize);
    return b;
}

// The variant (nco, ksplit) the library picks for a shape: the tuned code,
// else the default rule, with the generic kernel's tile and K split resolved.
// Under a fixed variant each output element's K sum runs in the same order
// whichever voxels share its tile, and every input index carries its sample,
// so a batch launched with the pick of its N = 1 shape gives each sample the
// N = 1 result bit for bit.
void conv_pick(const conv_args &a, int &nco, int &ksplit)
{
    const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
    const int ki = a.ksize * a.ksize * a.ksize * (a.cin / 16);
    nco = 0;
    ksplit = 0;
    const int code = dh_tune_lookup("conv3d", conv_key(a), -1);
    if (code >= 3000) { nco = code - 3000; ksplit = -3; }
    else if (code >= 1100) { nco = code - 1100; ksplit = -2; }
    else if (code >= 1000) { nco = code - 1000; ksplit = -1; }
    else if (code > 0) { nco = code / 100; ksplit = code % 100; }
    if (nco == 0 && ksplit == 0) {
        if (a.cout <= 16 && a.ksize == 1 && a.cb == 0 && a.up == 0) { nco = 1; ksplit = -3; }
        else if (a.cout <= 16 && a.ksize == 3 && nvox >= 32768) { nco = nvox >= (1 << 20) ? 4 : 8; ksplit = -3; }
        else if (a.ksize == 3 && nvox >= 4096 && a.cout_pad <= 64) { nco = 1; ksplit = -1; }
    }
    if (ksplit >= 0 && (nco <= 0 || ksplit <= 0)) conv_shape(nvox, a.cout_pad, ki, nco, ksplit);
}

void launch_conv3d(const conv_args &a, int nco, int ksplit, hipStream_t s)
{
    const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
    const int ki = a.ksize * a.ksize * a.ksize * (a.cin / 16);
    if (nco == 0 && ksplit == 0) {
        // tuned variant: 1000 + nco for the brick kernel (1100 + nco: two-row
        // bricks), 3000 + rows for the 16-channel kernels, else nco * 100 + ksplit
        const int code = dh_tune_lookup("conv3d", conv_key(a), -1);
        if (code >= 3000) { nco = code - 3000; ksplit = -3; }
        else if (code >= 1100) { nco = code - 1100; ksplit = -2; }
        else if (code >= 1000) { nco = code - 1000; ksplit = -1; }
        else if (code > 0) { nco = code / 100; ksplit = code % 100; }
    }
    if (nco == 0 && ksplit == 0) {
        // untuned shape: the variants the tuner picks for the surrogate's
        // layers timed over batches of points (16-output-channel kernels for
        // <= 16 outputs, four-row bricks from 2^20 voxels; four-row LDS bricks
        // for 3x3x3 layers of at most 64 outputs from 4096 voxels)
        if (a.cout <= 16 && a.ksize == 1 && a.cb == 0 && a.up == 0) { nco = 1; ksplit = -3; }
        else if (a.cout <= 16 && a.ksize == 3 && nvox >= 32768) { nco = nvox >= (1 << 20) ? 4 : 8; ksplit = -3; }
        else if (a.ksize == 3 && nvox >= 4096 && a.cout_pad <= 64) { nco = 1; ksplit = -1; }
    }
    int c = nco, k = ksplit;
    dh_timed t("nn_conv3d", s);
    // 16-channel brick kernel (ksplit = -3, nco = brick rows 4 or 8)
    if (ksplit == -3 && c == 1) {
        DH_REQUIRE(a.ksize == 1 && a.cout <= 16 && a.cb == 0 && a.up == 0,
                   "the 16-channel 1x1 kernel needs k = 1, at most 16 outputs and one source");
        const int64_t nvox = (int64_t)a.N * a.D * a.H * a.W;
        hipLaunchKernelGGL(k_conv1_c16, dim3((unsigned)((nvox + 255) / 256)), dim3(256), 0, s, a);
        DH_HIP(hipGetLastError());
        return;
    }
    if (ksplit == -3) {
        DH_REQUIRE(a.ksize == 3 && a.cout <= 16, "the 16-channel brick kernel is 3x3x3 with at most 16 outputs");
        DH_REQUIRE(c == 4 || c == 8, "16-channel brick kernel: rows must be 4 or 8");
        const int64_t items = (int64_t)a.N * ((a.D + 1) / 2) * ((a.H + c - 1) / c) * ((a.W + 31) / 32);
        if (c == 4) hipLaunchKernelGGL(k_conv3d_c16<4>, dim3((unsigned)items), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_conv3d_c16<8>, dim3((unsigned)items), dim3(256), 0, s, a);
        DH_HIP(hipGetLastError());
        return;
    }
    // brick kernel (ksplit = -1: four-row bricks, -2: two-row bricks)
    const bool brick = ksplit == -1 || ksplit == -2;
    if (brick) {
        DH_REQUIRE(a.ksize == 3, "the brick kernel is 3x3x3 only");
        if (c <= 0) c = a.cout_pad / 32;
        DH_REQUIRE((c == 1 || c == 2) && a.cout_pad % (32 * c) == 0, "brick kernel: nco must be 1 or 2");
        const int ty = ksplit == -2 ? 2 : 4;
        const int64_t items = (int64_t)a.N * ((a.D + 1) / 2) * ((a.H + ty - 1) / ty) * ((a.W + 31) / 32) *
                              (a.cout_pad / (32 * c));
        if (c == 1 && ty == 4)
            hipLaunchKernelGGL((k_conv3d_lds<1, 4>), dim3((unsigned)items), dim3(256), 0, s, a);
        else if (c == 2 && ty == 4)
            hipLaunchKernelGGL((k_conv3d_lds<2, 4>), dim3((unsigned)items), dim3(256), 0, s, a);
        else if (c == 1)
            hipLaunchKernelGGL((k_conv3d_lds<1, 2>), dim3((unsigned)items), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((k_conv3d_lds<2, 2>), dim3((unsigned)items), dim3(256), 0, s, a);
        DH_HIP(hipGetLastError());
        return;
    }
    if (c <= 0 || k <= 0) conv_shape(nvox, a.cout_pad, ki, c, k);
    DH_REQUIRE((a.cout_pad / 32) % c == 0, "output tile does not divide the padded channels");
    const int64_t items = ((nvox + 31) / 32) * (a.cout_pad / (32 * c));
    switch (c * 100 + k) {
    case 101: launch_conv<1, 1>(a, items, s); break;
    case 102: launch_conv<1, 2>(a, items, s); break;
    case 104: launch_conv<1, 4>(a, items, s); break;
    case 108: launch_conv<1, 8>(a, items, s); break;
    case 116: launch_conv<1, 16>(a, items, s); break;
    case 201: launch_conv<2, 1>(a, items, s); break;
    case 202: launch_conv<2, 2>(a, items, s); break;
    case 204: launch_conv<2, 4>(a, items, s); break;
    case 208: launch_conv<2, 8>(a, items, s); break;
    case 401: launch_conv<4, 1>(a, items, s); break;
    case 402: launch_conv<4, 2>(a, items, s); break;
    case 404: launch_conv<4, 4>(a, items, s); break;
    default: DH_REQUIRE(false, "unsupported (output tiles, K split) pair");
    }
    DH_HIP(hipGetLastError());
}

conv_args make_conv(const void *xa, int ca, int up, const void *xb, int cb, int n, int d, int h, int w,
                    const void *wt, const float *bias, int cout, int cout_pad, int ksize, int act,
                    float slope, void *y, float *y32)
{
    DH_REQUIRE(xa && wt && (y || y32), "null argument");
    DH_REQUIRE(ksize == 1 || ksize == 3, "kernel size must be 1 or 3");
    DH_REQUIRE(ca > 0 && ca % 16 == 0 && cb >= 0 && cb % 16 == 0 && (cb == 0 || xb),
               "input channels must be positive multiples of 16");
    DH_REQUIRE(up == 0 || up == 1, "up_a must be 0 or 1");
    DH_REQUIRE(!up || (d % 2 == 0 && h % 2 == 0 && w % 2 == 0), "upsampled input needs even sizes");
    DH_REQUIRE(cout > 0 && cout_pad >= cout && cout_pad % 32 == 0, "bad output channel padding");
    DH_REQUIRE(n > 0 && d > 0 && h > 0 && w > 0, "bad shape");
    DH_REQUIRE((int64_t)n * d * h * w * std::max(ca + cb, cout_pad) < ((int64_t)1 << 31) &&
                   (int64_t)27 * cout_pad * (ca + cb) < ((int64_t)1 << 31),
               "convolution tensors must hold fewer than 2^31 elements");
    conv_args a;
    a.xa = (const uint16_t *)xa; a.xb = (const uint16_t *)xb; a.ca = ca; a.cb = cb; a.up = up;
    a.w = (const uint16_t *)wt; a.bias = bias;
    a.y = (uint16_t *)y; a.y32 = y32;
    a.N = n; a.D = d; a.H = h; a.W = w; a.cin = ca + cb; a.cout = cout; a.cout_pad = cout_pad;
    a.ksize = ksize; a.act = act; a.slope = slope;
    return a;
}

}  // namespace

extern "C" int dh_nn_conv3d_dev(const void *xa, int32_t ca, int32_t up_a, const void *xb, int32_t cb,
                                int32_t n, int32_t d, int32_t h, int32_t w, const void *wt,
                                const float *bias, int32_t cout, int32_t cout_pad, int32_t ksize,
                                int32_t act, float slope, void *y, float *y32, int32_t nco,
                                int32_t ksplit, void *stream)
{
    return dh_guarded([&] {
        conv_args a = make_conv(xa, ca, up_a, xb, cb, n, d, h, w, wt, bias, cout, cout_pad, ksize, act, slope,
                                y, y32);
        launch_conv3d(a, nco, ksplit, (hipStream_t)stream);
    });
}

extern "C" int dh_nn_pool2_dev(const void *x, int32_t n, int32_t d, int32_t h, int32_t w, int32_t c,
                               void *y, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(x && y && n > 0 && c > 0, "bad argument");
        DH_REQUIRE(d > 1 && h > 1 && w > 1 && d % 2 == 0 && h % 2 == 0 && w % 2 == 0, "sizes must be even");
        launch_pool((const uint16_t *)x, n, d, h, w, c, (uint16_t *)y, (hipStream_t)stream);
    });
}

namespace {

// spacing channel: spacing / 2.5 mm
constexpr float kSpRef = 2.5f;

void launch_pack(int nz, int ny, int nx, const float *terma, const float *density, const float dir[3],
                 float spacing, uint16_t *x, float *tmax, hipStream_t s)
{
    const int64_t n = (int64_t)nz * ny * nx;
    const int cz = ((nz + 15) / 16) * 8, cy = ((ny + 15) / 16) * 8, cx = ((nx + 15) / 16) * 8;
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, tmax, 1, 0.0f);
    // a grid-stride loop over at most 256 workgroups: each ends in one atomic
    // on the same word, and thousands of them serialise for longer than the read
    hipLaunchKernelGGL(k_max_nonneg, dim3((unsigned)std::min<int64_t>((n / 4 + 255) / 256 + 1, 256)), dim3(256), 0, s,
                       terma, n, tmax);
    const int64_t items = (int64_t)cz * cy * cx;
    dh_timed t("sur_pack", s);
    hipLaunchKernelGGL(k_sur_pack, dim3((unsigned)std::min<int64_t>((items + 255) / 256, 8192)), dim3(256), 0, s,
                       terma, density, nz, ny, nx, cz, cy, cx, (const float *)tmax, dir[0], dir[1], dir[2],
                       spacing / kSpRef, x);
    DH_HIP(hipGetLastError());
}

// The packed inputs of n points of one grid and density (n <= 8) in one
// launch per step: both scalars of every point zeroed ({max TERMA, max
// dose} pairs in scal), the TERMA maxima, the packing.  Each point's result
// equals launch_pack's.
void launch_pack_batch(int nz, int ny, int nx, int n, const float *const *terma, const float *density,
                       const float (*dirs)[3], float spacing, uint16_t *x, int64_t x_stride, float *scal,
                       hipStream_t s)
{
    DH_REQUIRE(n >= 1 && n <= kPackBatchMax, "pack batch of 1..8 points");
    const int64_t nv = (int64_t)nz * ny * nx;
    const int cz = ((nz + 15) / 16) * 8, cy = ((ny + 15) / 16) * 8, cx = ((nx + 15) / 16) * 8;
    sur_pack_batch b;
    std::memset(&b, 0, sizeof(b));
    for (int i = 0; i < n; i++) {
        b.terma[i] = terma[i];
        b.tmax[i] = scal + 2 * i;
        b.x[i] = x + (size_t)i * x_stride;
        for (int c = 0; c < 3; c++) b.dir[i][c] = dirs[i][c];
    }
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, scal, 2 * n, 0.0f);
    hipLaunchKernelGGL(k_max_nonneg_batch, dim3((unsigned)std::min<int64_t>((nv / 4 + 255) / 256 + 1, 256), n),
                       dim3(256), 0, s, b, nv);
    const int64_t items = (int64_t)cz * cy * cx;
    dh_timed t("sur_pack", s);
    hipLaunchKernelGGL(k_sur_pack_batch, dim3((unsigned)std::min<int64_t>((items + 255) / 256, 8192), n), dim3(256), 0,
                       s, b, density, nz, ny, nx, cz, cy, cx, spacing / kSpRef);
    DH_HIP(hipGetLastError());
}

// zero_dmax false: the caller has zeroed *dmax (launch_pack_batch does)
void launch_finish(int nz, int ny, int nx, const float *y32, const float *tmax, float out_scale,
                   const float *terma, const float *electron, float mu_scale, float *dose, float *dmax,
                   hipStream_t s, bool zero_dmax = true)
{
    const int64_t n = (int64_t)nz * ny * nx;
    const int cy = ((ny + 15) / 16) * 8, cx = ((nx + 15) / 16) * 8;
    if (zero_dmax) hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, dmax, 1, 0.0f);
    dh_timed t("sur_finish", s);
    const int64_t ncell = (int64_t)((nz + 1) / 2) * ((ny + 1) / 2) * ((nx + 1) / 2);
    hipLaunchKernelGGL(k_sur_finish, dim3((unsigned)std::min<int64_t>((ncell + 255) / 256, 512)), dim3(256), 0, s,
                       y32, cy, cx, tmax, out_scale, terma, electron, mu_scale, nz, ny, nx, dose, dmax);
    DH_HIP(hipGetLastError());
}

void launch_pool(const uint16_t *x, int n, int d, int h, int w, int c, uint16_t *y, hipStream_t s)
{
    const int64_t total = (int64_t)n * (d / 2) * (h / 2) * (w / 2) * c;
    const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(k_pool2, dim3(grid), dim3(256), 0, s, x, n, d, h, w, c, y);
    DH_HIP(hipGetLastError());
}

}  // namespace

extern "C" int dh_sur_pack_dev(int32_t nz, int32_t ny, int32_t nx, const float *terma, const float *density,
                               const float dir[3], float spacing, void *x_out, float *tmax_out, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(terma && density && dir && x_out && tmax_out, "null argument");
        DH_REQUIRE(nz > 0 && ny > 0 && nx > 0 && spacing > 0.0f, "bad shape");
        launch_pack(nz, ny, nx, terma, density, dir, spacing, (uint16_t *)x_out, tmax_out, (hipStream_t)stream);
    });
}

extern "C" int dh_sur_finish_dev(int32_t nz, int32_t ny, int32_t nx, const float *y32, const float *tmax,
                                 float out_scale, const float *terma, const float *electron, float mu_scale,
                                 float *dose_out, float *dmax_out, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(y32 && tmax && terma && electron && dose_out && dmax_out, "null argument");
        DH_REQUIRE(nz > 0 && ny > 0 && nx > 0, "bad shape");
        launch_finish(nz, ny, nx, y32, tmax, out_scale, terma, electron, mu_scale, dose_out, dmax_out,
                      (hipStream_t)stream);
    });
}

// ---------------------------------------------------------------------------
// native surrogate network
// ---------------------------------------------------------------------------
struct dh_sur_layer {
    const void *w;           // bf16 [taps][cout_pad][ca + cb]
    const float *b;          // fp32 [cout]
    int ca, cb, up, cout, cout_pad, ksize, act, out32;
};

struct dh_sur_net {
    int gpu;
    int in_ch;
    int widths[4];
    float slope, out_scale;
    dh_sur_layer L[15];
};

namespace {

constexpr int kSurLayers = 15;
constexpr int kSurOut = 16;      // head output channels (0-7 used)
constexpr int kSurBatch = 8;     // control points per network batch
static_assert(kSurBatch <= kPackBatchMax, "a network batch must fit one packing launch");

// activation elements (bf16) of one forward over n samples of cz x cy x cx cells
int64_t sur_act_elems(const dh_sur_net *net, int n, int cz, int cy, int cx)
{
    const int64_t L0 = (int64_t)n * cz * cy * cx, L1 = L0 / 8, L2 = L0 / 64, L3 = L0 / 512;
    const int a = net->widths[0], b = net->widths[1], c = net->widths[2], d = net->widths[3];
    return L0 * 3 * a + L1 * (a + 3 * b) + L2 * (b + 3 * c) + L3 * (c + 2 * d);
}

// One layer over n samples with the variant picked for one sample, so every
// sample's output equals that of a forward over it alone.  The 16-output-
// channel brick kernels are the exception: their 4- and 8-row forms add
// every tile's taps and slices in the same order, so the batch takes the
// rows picked for the batch's own shape.
void sur_conv(const dh_sur_net *net, int li, const uint16_t *xa, const uint16_t *xb, int n, int d, int h, int w,
              void *y, float *y32, hipStream_t s)
{
    const dh_sur_layer &L = net->L[li];
    conv_args a = make_conv(xa, L.ca, L.up, xb, L.cb, 1, d, h, w, L.w, L.b, L.cout, L.cout_pad, L.ksize, L.act,
                            net->slope, y, y32);
    int nco, ksplit;
    conv_pick(a, nco, ksplit);
    a.N = n;
    if (n > 1 && ksplit == -3 && a.ksize == 3) {
        int nb, kb;
        conv_pick(a, nb, kb);
        if (kb == -3) nco = nb;
    }
    launch_conv3d(a, nco, ksplit, s);
}

// U-Net forward: x [n][cz][cy][cx][in_ch] bf16 -> y32 [n][cz][cy][cx][16] fp32
// With fin (the surrogate doses of the n points, finish arguments in fa):
// the output layer writes the doses directly (k_conv1_finish) when its
// picked variant is the 16-channel 1x1 kernel, whose arithmetic the fused
// kernel repeats; otherwise the layer writes y32 and k_sur_finish_batch
// follows.  Returns whether the doses were written.
struct sur_finish_args {
    float out_scale;
    int nz, ny, nx;
};
bool sur_forward(const dh_sur_net *net, const uint16_t *x, int n, int cz, int cy, int cx, float *y32,
                 uint16_t *ws, hipStream_t s, const sur_check_batch *fin = nullptr,
                 const sur_finish_args *fa = nullptr)
{
    const int64_t L0 = (int64_t)n * cz * cy * cx, L1 = L0 / 8, L2 = L0 / 64, L3 = L0 / 512;
    const int a = net->widths[0], b = net->widths[1], c = net->widths[2], d = net->widths[3];
    uint16_t *p = ws;
    auto take = [&](int64_t n) { uint16_t *q = p; p += n; return q; };
    uint16_t *A0 = take(L0 * a), *S0 = take(L0 * a), *B0 = take(L0 * a);
    uint16_t *P1 = take(L1 * a), *A1 = take(L1 * b), *S1 = take(L1 * b), *B1 = take(L1 * b);
    uint16_t *P2 = take(L2 * b), *A2 = take(L2 * c), *S2 = take(L2 * c), *B2 = take(L2 * c);
    uint16_t *P3 = take(L3 * c), *A3 = take(L3 * d), *S3 = take(L3 * d);
    const int z1 = cz / 2, y1 = cy / 2, x1 = cx / 2, z2 = cz / 4, y2 = cy / 4, x2 = cx / 4;
    const int z3 = cz / 8, y3 = cy / 8, x3 = cx / 8;
    sur_conv(net, 0, x, nullptr, n, cz, cy, cx, A0, nullptr, s);
    sur_conv(net, 1, A0, nullptr, n, cz, cy, cx, S0, nullptr, s);
    launch_pool(S0, n, cz, cy, cx, a, P1, s);
    sur_conv(net, 2, P1, nullptr, n, z1, y1, x1, A1, nullptr, s);
    sur_conv(net, 3, A1, nullptr, n, z1, y1, x1, S1, nullptr, s);
    launch_pool(S1, n, z1, y1, x1, b, P2, s);
    sur_conv(net, 4, P2, nullptr, n, z2, y2, x2, A2, nullptr, s);
    sur_conv(net, 5, A2, nullptr, n, z2, y2, x2, S2, nullptr, s);
    launch_pool(S2, n, z2, y2, x2, c, P3, s);
    sur_conv(net, 6, P3, nullptr, n, z3, y3, x3, A3, nullptr, s);
    sur_conv(net, 7, A3, nullptr, n, z3, y3, x3, S3, nullptr, s);
    sur_conv(net, 8, S3, S2, n, z2, y2, x2, B2, nullptr, s);
    sur_conv(net, 9, B2, nullptr, n, z2, y2, x2, A2, nullptr, s);
    sur_conv(net, 10, A2, S1, n, z1, y1, x1, B1, nullptr, s);
    sur_conv(net, 11, B1, nullptr, n, z1, y1, x1, A1, nullptr, s);
    sur_conv(net, 12, A1, S0, n, cz, cy, cx, B0, nullptr, s);
    sur_conv(net, 13, B0, nullptr, n, cz, cy, cx, A0, nullptr, s);
    if (fin) {
        const dh_sur_layer &L = net->L[14];
        conv_args ca = make_conv(A0, L.ca, L.up, nullptr, L.cb, 1, cz, cy, cx, L.w, L.b, L.cout, L.cout_pad, L.ksize,
                                 L.act, net->slope, nullptr, y32);
        int nco, ksplit;
        conv_pick(ca, nco, ksplit);
        const int64_t cells = (int64_t)cz * cy * cx;
        if (nco == 1 && ksplit == -3 && cells % 256 == 0) {
            ca.N = n;
            dh_timed t("sur_finish", s);
            hipLaunchKernelGGL(k_conv1_finish, dim3((unsigned)(n * cells / 256)), dim3(256), 0, s, ca, *fin,
                               fa->out_scale, fa->nz, fa->ny, fa->nx, cy, cx, cells);
            DH_HIP(hipGetLastError());
            return true;
        }
    }
    sur_conv(net, 14, A0, nullptr, n, cz, cy, cx, nullptr, y32, s);
    return false;
}

}  // namespace

extern "C" int dh_sur_net_create(int gpu_id, const int32_t widths[4], int32_t in_ch,
                                 const void *const *d_weights, const float *const *d_biases, float slope,
                                 float out_scale, dh_sur_net **out)
{
    return dh_guarded([&] {
        DH_REQUIRE(widths && d_weights && d_biases && out, "null argument");
        DH_REQUIRE(in_ch > 0 && in_ch % 16 == 0, "input channels must be a positive multiple of 16");
        for (int i = 0; i < 4; i++) DH_REQUIRE(widths[i] > 0 && widths[i] % 16 == 0, "widths must be multiples of 16");
        dh_device_enter(gpu_id);
        dh_sur_net *n = new dh_sur_net();
        n->gpu = gpu_id;
        n->in_ch = in_ch;
        for (int i = 0; i < 4; i++) n->widths[i] = widths[i];
        n->slope = slope;
        n->out_scale = out_scale;
        const int a = widths[0], b = widths[1], c = widths[2], d = widths[3];
        // (ca, cb, up, cout, ksize) of e0.0 e0.1 e1.0 e1.1 e2.0 e2.1 bt.0 bt.1
        // d2.0 d2.1 d1.0 d1.1 d0.0 d0.1 head
        const int tab[kSurLayers][5] = {{in_ch, 0, 0, a, 3}, {a, 0, 0, a, 3}, {a, 0, 0, b, 3}, {b, 0, 0, b, 3},
                                        {b, 0, 0, c, 3}, {c, 0, 0, c, 3}, {c, 0, 0, d, 3}, {d, 0, 0, d, 3},
                                        {d, c, 1, c, 3}, {c, 0, 0, c, 3}, {c, b, 1, b, 3}, {b, 0, 0, b, 3},
                                        {b, a, 1, a, 3}, {a, 0, 0, a, 3}, {a, 0, 0, kSurOut, 1}};
        for (int i = 0; i < kSurLayers; i++) {
            if (!d_weights[i] || !d_biases[i]) {
                delete n;
                DH_REQUIRE(false, "null layer weights");
            }
            dh_sur_layer &L = n->L[i];
            L.w = d_weights[i];
            L.b = d_biases[i];
            L.ca = tab[i][0]; L.cb = tab[i][1]; L.up = tab[i][2]; L.cout = tab[i][3]; L.ksize = tab[i][4];
            L.cout_pad = (L.cout + 31) / 32 * 32;
            L.act = i < kSurLayers - 1;
            L.out32 = i == kSurLayers - 1;
        }
        *out = n;
    });
}

int dh_sur_net_gpu(const dh_sur_net *net) { return net->gpu; }

extern "C" int dh_sur_net_destroy(dh_sur_net *net)
{
    delete net;
    return DH_OK;
}

extern "C" int dh_sur_net_forward_dev(const dh_sur_net *net, const void *d_x, int32_t cz, int32_t cy, int32_t cx,
                                      float *d_y32, void *stream)
{
    return dh_guarded([&] {
        DH_REQUIRE(net && d_x && d_y32, "null argument");
        DH_REQUIRE(cz > 0 && cy > 0 && cx > 0 && cz % 8 == 0 && cy % 8 == 0 && cx % 8 == 0,
                   "cell counts must be positive multiples of 8");
        dh_device_enter(net->gpu);
        uint16_t *ws = (uint16_t *)dh_scratch(net->gpu, DH_SLOT_NN_ACT,
                                              sizeof(uint16_t) * (size_t)sur_act_elems(net, 1, cz, cy, cx));
        sur_forward(net, (const uint16_t *)d_x, 1, cz, cy, cx, d_y32, ws, (hipStream_t)stream);
    });
}

namespace {
// offset (ints) of the last point's work buffer in the check slot after the
// last dh_launch_sur_dose_batch of this thread, per device (dh_sur_counts)
thread_local int64_t t_last_counts[64] = {};

// Lattice check and plain fallback of the n (<= 8) points of b, whose
// surrogate doses and dose maxima are final and whose counters are zero:
// one launch per step for all points (pack, lattice, lattice cones, flags,
// dilation, fallback list, fallback cones), the cone launches spreading all
// points' lists over the chip.
void sur_check_multi(int dev, const dh_ccc_args *cas, int n, const sur_check_batch &b, const float *density,
                     int stride, float tol, int dilate, hipStream_t s)
{
    const dh_geom &g = cas[0].g;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const int lz = (nz + stride - 1) / stride, ly = (ny + stride - 1) / stride, lx = (nx + stride - 1) / stride;
    const int64_t nc = (int64_t)lz * ly * lx, nv = (int64_t)nz * ny * nx;
    dh_ccc_multi m;
    dh_ccc_multi_init(m, cas, n);
    const size_t trb = dh_ccc_multi_tr_bytes(cas[0]), pb = dh_ccc_multi_part_bytes(cas[0]);
    char *buf = (char *)dh_scratch(dev, DH_SLOT_CHK_TR, (trb + pb) * (size_t)n);
    for (int q = 0; q < n; q++) {
        m.tr[q] = buf + (trb + pb) * (size_t)q;
        m.part[q] = (float *)(buf + (trb + pb) * (size_t)q + trb);
        m.terma[q] = b.terma[q];
        m.electron[q] = b.electron[q];
    }
    dh_launch_pack_tr_multi(m, density, s);
    const unsigned cgrid = (unsigned)((nc + 255) / 256);
    {
        dh_timed t("sur_lattice", s);
        hipLaunchKernelGGL(k_sur_lattice, dim3(cgrid, n), dim3(256), 0, s, b, nz, ny, nx, stride, lz, ly, lx);
    }
    for (int q = 0; q < n; q++) {
        m.list[q] = b.work[q] + 4 + 3 * nc;
        m.counter[q] = b.work[q];
        m.out[q] = b.plain[q];
    }
    dh_launch_ccc_multi("ccc_list", m, nc, s);
    hipLaunchKernelGGL(k_sur_flag, dim3(cgrid, n), dim3(256), 0, s, b, tol, ny, nx, stride, lz, ly, lx);
    {
        dh_timed t("sur_fallback", s);
        hipLaunchKernelGGL(k_sur_dilate, dim3(cgrid, n), dim3(256), 0, s, b, lz, ly, lx, dilate & 3);
        const unsigned bricks = (unsigned)((nx + kFbBx - 1) / kFbBx) * (unsigned)((ny + kFbBy - 1) / kFbBy) *
                                (unsigned)((nz + kFbBz - 1) / kFbBz);
        hipLaunchKernelGGL(k_sur_fallback, dim3(bricks, n), dim3(256), 0, s, b, (dilate & 4) ? 1 : 0, nz, ny, nx,
                           stride, lz, ly, lx);
    }
    for (int q = 0; q < n; q++) {
        m.list[q] = b.work[q] + 4 + 4 * nc;
        m.counter[q] = b.work[q] + 1;
        m.out[q] = b.dose[q];
    }
    dh_launch_ccc_multi("ccc_list", m, nv, s);
    DH_HIP(hipGetLastError());
}
}  // namespace

// Checked surrogate doses of ncp control points on one grid (one density,
// per-point TERMA and electron volumes): each point is packed into its slot
// of one batch and the network runs once per batch of up to kSurBatch
// points; then one finish launch and one launch per check step serve all
// points of the batch (sur_check_multi), each point with buffers of its own.
// Every layer uses the variant of its one-sample shape and every voxel the
// same arithmetic whatever the batch, so each dose equals the single-point
// result bit for bit.
void dh_launch_sur_dose_batch(const dh_sur_net *net, const dh_ccc_args *cas, int ncp,
                              const float *const *terma, const float *density, const float *const *electron,
                              const float (*dirs)[3], int stride, float tol, int dilate, float *const *dose,
                              hipStream_t s)
{
    DH_REQUIRE(net->in_ch == 32, "the surrogate input layout has 32 channels");
    DH_REQUIRE(ncp >= 1, "no control points");
    const dh_geom &g = cas[0].g;
    const int nz = g.nz, ny = g.ny, nx = g.nx;
    const int cz = ((nz + 15) / 16) * 8, cy = ((ny + 15) / 16) * 8, cx = ((nx + 15) / 16) * 8;
    const int64_t cells = (int64_t)cz * cy * cx;
    const int nb = std::min(ncp, kSurBatch);
    const int dev = net->gpu;
    DH_REQUIRE(dev >= 0 && dev < 64, "device id out of range");
    uint16_t *x = (uint16_t *)dh_scratch(dev, DH_SLOT_NN_X, sizeof(uint16_t) * (size_t)(nb * cells * net->in_ch));
    uint16_t *ws = (uint16_t *)dh_scratch(dev, DH_SLOT_NN_ACT,
                                          sizeof(uint16_t) * (size_t)sur_act_elems(net, nb, cz, cy, cx));
    float *y32 = (float *)dh_scratch(dev, DH_SLOT_NN_Y, sizeof(float) * (size_t)(nb * cells * kSurOut));
    float *scal = (float *)dh_scratch(dev, DH_SLOT_NN_SCAL, 256);     // {max TERMA, max dose} per slot
    const int64_t wl = dh_sur_check_work_len(nz, ny, nx, stride);
    const int64_t nv = (int64_t)nz * ny * nx;
    int *work = (int *)dh_scratch(dev, DH_SLOT_NN_WORK, sizeof(int) * (size_t)(wl * nb));
    float *plain = (float *)dh_scratch(dev, DH_SLOT_NN_PLAIN, sizeof(float) * (size_t)(nv * nb));
    for (int c0 = 0; c0 < ncp; c0 += nb) {
        const int n = std::min(nb, ncp - c0);
        launch_pack_batch(nz, ny, nx, n, terma + c0, density, dirs + c0, g.s, x, cells * net->in_ch, scal, s);
        sur_check_batch b;
        std::memset(&b, 0, sizeof(b));
        for (int i = 0; i < n; i++) {
            const dh_ccc_args &ca = cas[c0 + i];
            b.y32[i] = y32 + (size_t)i * cells * kSurOut;
            b.tmax[i] = scal + 2 * i;
            b.mu_scale[i] = ca.mu_cal * ca.mu;
            b.terma[i] = terma[c0 + i];
            b.electron[i] = electron[c0 + i];
            b.dose[i] = dose[c0 + i];
            b.dmax[i] = scal + 2 * i + 1;
            b.plain[i] = plain + (size_t)i * nv;
            b.work[i] = work + (size_t)i * wl;
        }
        const sur_finish_args fa{net->out_scale, nz, ny, nx};
        if (!sur_forward(net, x, n, cz, cy, cx, y32, ws, s, &b, &fa)) {
            dh_timed t("sur_finish", s);
            // about 2048 workgroups over the batch (at least 512 per point)
            const int fgrid = std::min((nz + 1) / 2 * ((ny + 1) / 2), std::max(512, 2048 / n));
            hipLaunchKernelGGL(k_sur_finish_batch, dim3((unsigned)fgrid, n), dim3(256), 0, s, b, cy, cx,
                               net->out_scale, nz, ny, nx);
            DH_HIP(hipGetLastError());
        }
        sur_check_multi(dev, cas + c0, n, b, density, stride, tol, dilate, s);
        t_last_counts[dev] = (int64_t)(n - 1) * wl;
    }
}

void dh_launch_sur_check(int dev, const dh_ccc_args &ca, const float *terma, const float *density,
                         const float *electron, float *dose, const float *dmax, int stride, float tol, int dilate,
                         int *work, float *plain, hipStream_t s)
{
    sur_check_batch b;
    std::memset(&b, 0, sizeof(b));
    b.terma[0] = terma;
    b.electron[0] = electron;
    b.dose[0] = dose;
    b.dmax[0] = const_cast<float *>(dmax);     // read only by the check kernels
    b.plain[0] = plain;
    b.work[0] = work;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, work, 4, 0);
    sur_check_multi(dev, &ca, 1, b, density, stride, tol, dilate, s);
}

void dh_launch_sur_dose(const dh_sur_net *net, const dh_ccc_args &ca, const float *terma,
                        const float *density, const float *electron, const float dir[3],
                        int stride, float tol, int dilate, float *dose, hipStream_t s)
{
    const float d[1][3] = {{dir[0], dir[1], dir[2]}};
    dh_launch_sur_dose_batch(net, &ca, 1, &terma, density, &electron, d, stride, tol, dilate, &dose, s);
}

extern "C" int dh_sur_counts(int gpu_id, int32_t *lattice, int32_t *fallback)
{
    return dh_guarded([&] {
        DH_REQUIRE(lattice && fallback, "null argument");
        dh_device_enter(gpu_id);
        DH_REQUIRE(gpu_id >= 0 && gpu_id < 64, "device id out of range");
        const int64_t off = t_last_counts[gpu_id];
        int *work = (int *)dh_scratch(gpu_id, DH_SLOT_NN_WORK, sizeof(int) * (size_t)(off + 4)) + off;
        int h[2] = {0, 0};
        DH_HIP(hipDeviceSynchronize());
        DH_HIP(hipMemcpy(h, work, sizeof(h), hipMemcpyDeviceToHost));
        *lattice = h[0];
        *fallback = h[1];
    });
}

int64_t dh_sur_check_work_len(int nz, int ny, int nx, int stride)
{
    const int64_t nc = (int64_t)((nz + stride - 1) / stride) * ((ny + stride - 1) / stride) *
                       ((nx + stride - 1) / stride);
    return 4 + 4 * nc + (int64_t)nz * ny * nx;
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

namespace {

__global__ void k_fill_bf16_hash(uint16_t *p, int64_t n, uint32_t seed, float scale)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t h = (uint32_t)i * 2654435761u ^ seed;
        h ^= h >> 15; h *= 2246822519u; h ^= h >> 13;
        p[i] = f2bf(((float)(h & 0xffffu) / 65536.0f - 0.5f) * scale);
    }
}

}  // namespace

// Tunes the convolution variant of one layer shape on synthetic data: every
// (output tiles, K split) pair and the brick kernel; a variant is admitted
// when its fp32 output matches the first variant's to 1e-5 of the largest
// magnitude (the variants differ only in summation order); the fastest is
// stored for the shape.  codes[] / ms[] receive up to 24 tried variants.
extern "C" int dh_autotune_conv3d(int gpu_id, int32_t n, int32_t d, int32_t h, int32_t w, int32_t ca, int32_t cb,
                                  int32_t up, int32_t cout, int32_t ksize, int32_t reps, int32_t *best_code,
                                  int32_t *codes, float *ms, int32_t *n_tried, void *stream)
{
    return dh_guarded([&] {
        dh_device_enter(gpu_id);
        hipStream_t s = (hipStream_t)stream;
        const int cout_pad = (cout + 31) / 32 * 32;
        const int cin = ca + cb;
        const int64_t nvox = (int64_t)n * d * h * w;
        const int64_t na = (int64_t)n * (d >> up) * (h >> up) * (w >> up) * ca, nb = nvox * cb;
        const int64_t nw = (int64_t)ksize * ksize * ksize * cout_pad * cin;
        uint16_t *xa = (uint16_t *)dh_scratch(gpu_id, DH_SLOT_IN0, sizeof(uint16_t) * (size_t)na);
        uint16_t *xb = cb ? (uint16_t *)dh_scratch(gpu_id, DH_SLOT_IN1, sizeof(uint16_t) * (size_t)nb) : nullptr;
        uint16_t *wt = (uint16_t *)dh_scratch(gpu_id, DH_SLOT_TMP0, sizeof(uint16_t) * (size_t)nw);
        float *ref = (float *)dh_scratch(gpu_id, DH_SLOT_OUT0, sizeof(float) * (size_t)(nvox * cout));
        float *out = (float *)dh_scratch(gpu_id, DH_SLOT_OUT1, sizeof(float) * (size_t)(nvox * cout));
        hipLaunchKernelGGL(k_fill_bf16_hash, dim3(1024), dim3(256), 0, s, xa, na, 17u, 2.0f);
        if (cb) hipLaunchKernelGGL(k_fill_bf16_hash, dim3(1024), dim3(256), 0, s, xb, nb, 29u, 2.0f);
        hipLaunchKernelGGL(k_fill_bf16_hash, dim3(1024), dim3(256), 0, s, wt, nw, 41u, 0.2f);
        std::vector<int> cand;
        if (ksize == 3)
            for (int c : {1, 2})
                if (cout_pad % (32 * c) == 0) {
                    cand.push_back(1000 + c);
                    cand.push_back(1100 + c);
                }
        if (ksize == 3 && cout <= 16) {
            cand.push_back(3004);
            cand.push_back(3008);
        }
        if (ksize == 1 && cout <= 16 && cb == 0 && up == 0) cand.push_back(3001);
        for (int c : {1, 2, 4})
            for (int k : {1, 2, 4, 8, 16}) {
                if ((cout_pad / 32) % c || c * k > 16 || (c == 2 && k == 16) || (c == 4 && k > 4)) continue;
                cand.push_back(c * 100 + k);
            }
        std::vector<float> host_ref, host_out;
        int best = -1, tried = 0;
        float best_ms = 0.0f;
        for (size_t q = 0; q < cand.size() && tried < 24; q++) {
            const int code = cand[q];
            const int c = code >= 3000 ? code - 3000 : code >= 1100 ? code - 1100 : code >= 1000 ? code - 1000 : code / 100;
            const int k = code >= 3000 ? -3 : code >= 1100 ? -2 : code >= 1000 ? -1 : code % 100;
            conv_args a = make_conv(xa, ca, up, xb, cb, n, d, h, w, wt, nullptr, cout, cout_pad, ksize, 0, 0.1f,
                                    nullptr, tried == 0 ? ref : out);
            const float t = dh_time_median([&] { launch_conv3d(a, c, k, s); }, reps, s);
            bool ok = true;
            if (tried > 0) {
                host_ref.resize((size_t)(nvox * cout));
                host_out.resize((size_t)(nvox * cout));
                DH_HIP(hipMemcpyAsync(host_ref.data(), ref, sizeof(float) * host_ref.size(), hipMemcpyDeviceToHost, s));
                DH_HIP(hipMemcpyAsync(host_out.data(), out, sizeof(float) * host_out.size(), hipMemcpyDeviceToHost, s));
                DH_HIP(hipStreamSynchronize(s));
                float mx = 0.0f, md = 0.0f;
                for (size_t i = 0; i < host_ref.size(); i++) {
                    mx = std::max(mx, std::fabs(host_ref[i]));
                    md = std::max(md, std::fabs(host_ref[i] - host_out[i]));
                }
                ok = md <= 1e-5f * mx;
            }
            if (codes) codes[tried] = ok ? code : -code;
            if (ms) ms[tried] = t;
            tried++;
            if (ok && (best < 0 || t < best_ms)) { best = code; best_ms = t; }
        }
        DH_REQUIRE(best > 0, "no convolution variant ran");
        conv_args a = make_conv(xa, ca, up, xb, cb, n, d, h, w, wt, nullptr, cout, cout_pad, ksize, 0, 0.1f, nullptr,
                                out);
        dh_tune_store("conv3d", conv_key(a), best, best_ms);
        if (best_code) *best_code = best;
        if (n_tried) *n_tried = tried;
    });
}
