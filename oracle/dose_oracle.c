/*
 * dose_oracle.c -- FP64 host oracle for the DoseCUDA dose engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped engine (dosecuda_amd/)
 * links, loads or calls this file; only tests/, the smoke check and the
 * cpu_baseline leg of bench.py use it, to check and to time the engine's
 * results, never as a result of the engine.
 *
 * It restates, per voxel, the formulae of the reference kernels:
 *   rayTraceKernel      DoseCUDA/dose_kernels/CudaClasses.cu:180-228
 *   smoothRayKernel     DoseCUDA/dose_kernels/IMPTClasses.cu:139-202
 *   pencilBeamKernel    DoseCUDA/dose_kernels/IMPTClasses.cu:204-271
 *   termaKernel         DoseCUDA/dose_kernels/IMRTClasses.cu:209-250
 *   cccKernel           DoseCUDA/dose_kernels/IMRTClasses.cu:252-350
 * plus the boundary set-up of DoseCUDA/dose_kernels/dosemodule.cu:104-557.
 *
 * Precision policy.  Accumulations, sampled values, transcendental functions
 * and every continuous quantity are FP64.  The *decisions* of the reference
 * (which samples exist on a ray, where a ray leaves the volume, how many
 * steps a cone takes, which neighbour a rounded index picks, the range cut)
 * are made with the same FP32 operation sequence the GPU engine uses, so
 * that a one-ulp difference can never add or drop a whole sample.  The FP32
 * sequences are written without fused multiply-adds except where the
 * reference itself calls fmaf(); this file is compiled with
 * -ffp-contract=off so the compiler does not fuse on its own.
 *
 * Texture sampling (DoseCUDA/dose_kernels/TextureClasses.cu:14-41) is
 * restated as exact trilinear interpolation with border addressing and
 * unnormalised coordinates; the reduced-precision fixed-point weights of
 * NVIDIA texture hardware are not emulated.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI_F 3.141592654f /* CUDART_PI_F */
#define ORC_PI 3.14159265358979323846
#define ORC_LUT_LENGTH 400    /* IMPTClasses.cuh:6 */
#define ORC_AIR_DENSITY 0.0012f /* IMRTClasses.cuh:6 */

typedef struct {
    int nx, ny, nz;          /* CudaClasses.cu:168-178: i = x fastest */
    float s;                 /* spacing[0] (dosemodule.cu:146) */
    float inv_s;             /* 1.0f / s  (CudaClasses.cuh:101) */
    float iso[3];            /* iso - origin (dosemodule.cu:160-164) */
    float src[3];            /* CudaClasses.cu:17-51 */
    float singa, cosga, sinta, costa, sinca, cosca;
} orc_geom;

/* ------------------------------------------------------------------ */
/* Set-up: dosemodule.cu:148-172 and CudaClasses.cu:17-51,            */
/* IMRTClasses.cu:30-39.                                              */
/* ------------------------------------------------------------------ */
void orc_geom_init(orc_geom *g, int nz, int ny, int nx, float spacing,
                   const float origin[3], const float iso[3],
                   double gantry, double couch, double collimator,
                   float src_dist)
{
    g->nx = nx; g->ny = ny; g->nz = nz;
    g->s = spacing;
    g->inv_s = 1.0f / spacing;
    for (int a = 0; a < 3; a++) g->iso[a] = iso[a] - origin[a];
    float adj = fmodf((float)(gantry + 180.0), 360.0f);
    float ga = adj * ORC_PI_F / 180.0f;
    float ta = (float)couch * ORC_PI_F / 180.0f;
    float ca = (float)collimator * ORC_PI_F / 180.0f;
    g->singa = sinf(ga); g->cosga = cosf(ga);
    g->sinta = sinf(ta); g->costa = cosf(ta);
    g->sinca = sinf(ca); g->cosca = cosf(ca);
    float xg = -src_dist * g->singa;
    float yg = src_dist * g->cosga;
    g->src[0] = xg * g->costa;
    g->src[1] = yg;
    g->src[2] = -xg * g->sinta;
}

/* ---------------- FP32 geometry (decision arithmetic) --------------- */

static void vox_xyz_f(const orc_geom *g, int i, int j, int k, float p[3])
{   /* CudaClasses.cuh:106-110 */
    p[0] = (float)i * g->s - g->iso[0];
    p[1] = (float)j * g->s - g->iso[1];
    p[2] = (float)k * g->s - g->iso[2];
}

static void tex_f(const orc_geom *g, const float q[3], float t[3])
{   /* CudaClasses.cuh:100-104 */
    for (int a = 0; a < 3; a++) t[a] = fmaf(g->inv_s, q[a] + g->iso[a], 0.5f);
}

static int tex_inside_f(const orc_geom *g, const float t[3])
{   /* CudaClasses.cuh:90-94 */
    return t[0] >= 0.0f && t[0] < (float)g->nx &&
           t[1] >= 0.0f && t[1] < (float)g->ny &&
           t[2] >= 0.0f && t[2] < (float)g->nz;
}

/* unit vector from p to the source, CudaClasses.cuh:41-52.  The norm is the
 * correctly rounded 1/sqrt of the FP32 sum of squares (rnorm3df's own ulp
 * behaviour cannot be pinned here). */
static void uvec_f(const orc_geom *g, const float p[3], float u[3])
{
    float x = g->src[0] - p[0], y = g->src[1] - p[1], z = g->src[2] - p[2];
    float ss = x * x + y * y + z * z;
    float r = (float)(1.0 / sqrt((double)ss));
    u[0] = x * r; u[1] = y * r; u[2] = z * r;
}

/* Image -> head including the collimator step (IMRTClasses.cu:41-79); with
 * sinca = 0, cosca = 1 the collimator step is the identity and this is
 * CudaClasses.cu:65-95. */
static void img_to_head_f(const orc_geom *g, const float p[3], float h[3])
{
    float sx = -g->sinta, cx = g->costa;
    float xt = p[0] * cx + p[2] * sx;
    float yt = p[1];
    float zt = -p[0] * sx + p[2] * cx;
    sx = -g->singa; cx = g->cosga;
    float xg = xt * cx - yt * sx;
    float yg = xt * sx + yt * cx;
    float zg = zt;
    sx = -g->sinca; cx = g->cosca;
    float xc = xg * cx + zg * sx;
    float yc = yg;
    float zc = -xg * sx + zg * cx;
    h[0] = -xc; h[1] = zc; h[2] = yc;
}

/* Head -> image (IMRTClasses.cu:81-118; CudaClasses.cu:97-126). */
static void head_to_img_f(const orc_geom *g, const float h[3], float p[3])
{
    float xz = -h[0], yz = h[2], zz = h[1];
    float sx = g->sinca, cx = g->cosca;
    float xc = xz * cx + zz * sx;
    float yc = yz;
    float zc = -xz * sx + zz * cx;
    sx = g->singa; cx = g->cosga;
    float xg = xc * cx - yc * sx;
    float yg = xc * sx + yc * cx;
    float zg = zc;
    sx = g->sinta; cx = g->costa;
    p[0] = xg * cx + zg * sx;
    p[1] = yg;
    p[2] = -xg * sx + zg * cx;
}

/* ---------------- FP64 continuous geometry -------------------------- */

static void img_to_head_d(const orc_geom *g, const double p[3], double h[3])
{
    double sx = -(double)g->sinta, cx = g->costa;
    double xt = p[0] * cx + p[2] * sx, yt = p[1], zt = -p[0] * sx + p[2] * cx;
    sx = -(double)g->singa; cx = g->cosga;
    double xg = xt * cx - yt * sx, yg = xt * sx + yt * cx, zg = zt;
    sx = -(double)g->sinca; cx = g->cosca;
    double xc = xg * cx + zg * sx, yc = yg, zc = -xg * sx + zg * cx;
    h[0] = -xc; h[1] = zc; h[2] = yc;
}

static void head_to_img_d(const orc_geom *g, const double h[3], double p[3])
{
    double xz = -h[0], yz = h[2], zz = h[1];
    double sx = g->sinca, cx = g->cosca;
    double xc = xz * cx + zz * sx, yc = yz, zc = -xz * sx + zz * cx;
    sx = g->singa; cx = g->cosga;
    double xg = xc * cx - yc * sx, yg = xc * sx + yc * cx, zg = zc;
    sx = g->sinta; cx = g->costa;
    p[0] = xg * cx + zg * sx; p[1] = yg; p[2] = -xg * sx + zg * cx;
}

/* ---------------- trilinear sampling with border -------------------- */

static double vol_at(const float *v, int nx, int ny, int nz,
                     long i, long j, long k, double border)
{
    if (i < 0 || j < 0 || k < 0 || i >= nx || j >= ny || k >= nz) return border;
    return (double)v[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)];
}

/* TextureClasses.cu:22-38: border addressing, linear filter, unnormalised
 * coordinates; the texel centre of voxel i sits at t = i + 0.5. */
double orc_sample(const float *v, int nx, int ny, int nz,
                  double tx, double ty, double tz, double border)
{
    double u = tx - 0.5, w = ty - 0.5, z = tz - 0.5;
    double fu = floor(u), fw = floor(w), fz = floor(z);
    long i0 = (long)fu, j0 = (long)fw, k0 = (long)fz;
    double ax = u - fu, ay = w - fw, az = z - fz;
    double acc = 0.0;
    for (int dk = 0; dk < 2; dk++)
        for (int dj = 0; dj < 2; dj++)
            for (int di = 0; di < 2; di++) {
                double wt = (di ? ax : 1.0 - ax) * (dj ? ay : 1.0 - ay) * (dk ? az : 1.0 - az);
                if (wt == 0.0) continue;
                acc += wt * vol_at(v, nx, ny, nz, i0 + di, j0 + dj, k0 + dk, border);
            }
    return acc;
}

/* ------------------------------------------------------------------ */
/* rayTraceKernel, CudaClasses.cu:180-228.  Returns the WET in cm     */
/* and (optionally) the number of samples on each ray.                */
/* ------------------------------------------------------------------ */
/* `mask` (optional): only voxels with mask[v] != 0 are traced (the others
 * get 0); the CPU baseline times a slab of voxels this way. */
void orc_raytrace_m(const orc_geom *g, const float *rho, double border,
                    double *wet_out, int32_t *nsamp_out, const uint8_t *mask)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    #pragma omp parallel for schedule(dynamic, 4) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                if (mask && !mask[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k)]) {
                    size_t v0 = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
                    wet_out[v0] = 0.0;
                    if (nsamp_out) nsamp_out[v0] = 0;
                    continue;
                }
                float p[3], u[3], q[3], t[3];
                vox_xyz_f(g, i, j, k, p);
                uvec_f(g, p, u);
                float L = 0.0f;
                double sum = 0.0;
                int n = 0;
                for (;;) {
                    for (int a = 0; a < 3; a++) q[a] = fmaf(u[a], L, p[a]);
                    tex_f(g, q, t);
                    if (!tex_inside_f(g, t)) break;
                    double d = orc_sample(rho, nx, ny, nz, t[0], t[1], t[2], border);
                    sum += d > 0.0 ? d : 0.0;
                    n++;
                    L += 1.0f;
                }
                size_t v = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
                wet_out[v] = -0.05 + 0.1 * sum;
                if (nsamp_out) nsamp_out[v] = n;
            }
}

void orc_raytrace(const orc_geom *g, const float *rho, double border,
                  double *wet_out, int32_t *nsamp_out)
{
    orc_raytrace_m(g, rho, border, wet_out, nsamp_out, NULL);
}

/* ------------------------------------------------------------------ */
/* smoothRayKernel, IMPTClasses.cu:139-202.  `wet` is the FP32 WET    */
/* (cm) the smoothing reads; the neighbour set is decided in FP32.    */
/* ------------------------------------------------------------------ */
void orc_smooth(const orc_geom *g, const float *wet, double *out)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    float sn[6], cs[6];
    for (int a = 0; a < 6; a++) {
        float ang = (float)a * ORC_PI_F / 3.0f;
        sn[a] = sinf(ang); cs[a] = cosf(ang);
    }
    #pragma omp parallel for schedule(static) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t v = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
                float p[3], h[3], hc[3], q[3];
                vox_xyz_f(g, i, j, k, p);
                img_to_head_f(g, p, h);
                float c = wet[v];
                double sum = (double)c;
                int n = 1;
                for (int a = 0; a < 6; a++) {
                    for (float dr = 1.0f; dr < c * 10.0f && dr < 10.0f; dr += 1.0f) {
                        hc[0] = h[0] + dr * cs[a];
                        hc[1] = h[1] + dr * sn[a];
                        hc[2] = h[2];
                        head_to_img_f(g, hc, q);
                        unsigned ii = (unsigned)(int)roundf((q[0] + g->iso[0]) / g->s);
                        unsigned jj = (unsigned)(int)roundf((q[1] + g->iso[1]) / g->s);
                        unsigned kk = (unsigned)(int)roundf((q[2] + g->iso[2]) / g->s);
                        if (ii < (unsigned)nx && jj < (unsigned)ny && kk < (unsigned)nz) {
                            sum += (double)wet[(size_t)ii + (size_t)nx * ((size_t)jj + (size_t)ny * (size_t)kk)];
                            n++;
                        }
                    }
                }
                out[v] = sum / (double)n;
            }
}

/* ------------------------------------------------------------------ */
/* Layers: dosemodule.cu:204-225 (spot array) + IMPTClasses.cu:29-67. */
/* Spots are sorted by (int) energy id with a stable sort; layers are */
/* the contiguous runs counted from energy id 0 upward, exactly as    */
/* importLayers does (spots with an id outside [0, n_energies) or     */
/* after a gap-breaking negative id are not assigned to any layer).   */
/* ------------------------------------------------------------------ */
typedef struct { float x, y, mu; int id; } orc_spot;
typedef struct { int start, n, id; float r80, energy; } orc_layer;

int orc_make_layers(const float *spots_rows, int n_spots, int n_energies,
                    const float *dvp /* [n_e,5] */, orc_spot *spots_out,
                    orc_layer *layers_out)
{
    /* read rows (dosemodule.cu:211-222), then a stable insertion sort by id */
    for (int i = 0; i < n_spots; i++) {
        orc_spot s;
        s.x = spots_rows[4 * i];
        s.y = spots_rows[4 * i + 1];
        s.mu = spots_rows[4 * i + 2];
        s.id = (int)spots_rows[4 * i + 3];
        int j = i;
        while (j > 0 && spots_out[j - 1].id > s.id) { spots_out[j] = spots_out[j - 1]; j--; }
        spots_out[j] = s;
    }
    int n_layers = 0, start = 0;
    for (int e = 0; e < n_energies; e++) {
        int end = start;
        while (end < n_spots && spots_out[end].id == e) end++;
        int cnt = end - start;
        if (!cnt) continue;
        orc_layer *L = &layers_out[n_layers++];
        L->start = start; L->n = cnt; L->id = e;
        L->r80 = dvp[5 * e + 1];
        L->energy = dvp[5 * e + 0];
        start += cnt;
    }
    return n_layers;
}

/* lowerBound, CudaClasses.cuh:130-144 */
static int lower_bound_f(const float *d, int len, float key)
{
    int l = 0, r = len;
    while (l < r) {
        int m = (l + r) / 2;
        if (key <= d[m]) r = m; else l = m + 1;
    }
    return l;
}

/* ------------------------------------------------------------------ */
/* pencilBeamKernel, IMPTClasses.cu:204-271 (+ helpers 69-133).       */
/* lut_* are [n_energies][lut_stride] with the first 400 used.        */
/* Returns the number of voxel-spot pairs that passed the range cut.  */
/* ------------------------------------------------------------------ */
int64_t orc_pencil(const orc_geom *g, float vsadx, float vsady,
                   const float *wet, const orc_spot *spots,
                   const orc_layer *layers, int n_layers,
                   const float *dvp, const float *lut_d, const float *lut_s,
                   const float *lut_i, int lut_stride, double *dose_out)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int64_t pairs = 0;
    #pragma omp parallel for schedule(dynamic, 2) collapse(2) reduction(+:pairs)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t v = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
                float wf = wet[v] * 10.0f;
                float pf[3], hf[3];
                vox_xyz_f(g, i, j, k, pf);
                double p[3] = { pf[0], pf[1], pf[2] }, h[3];
                img_to_head_d(g, p, h);
                double dx = g->src[0] - p[0], dy = g->src[1] - p[1], dz = g->src[2] - p[2];
                double dist = sqrt(dx * dx + dy * dy + dz * dz);
                (void)hf;
                double total = 0.0;
                for (int l = 0; l < n_layers; l++) {
                    const orc_layer *L = &layers[l];
                    float r80f = L->r80;
                    if (wf > 1.1f * r80f) continue;
                    double w = (double)wf, r80 = (double)r80f;
                    const float *dd = lut_d + (size_t)L->id * lut_stride;
                    const float *ss = lut_s + (size_t)L->id * lut_stride;
                    const float *ii = lut_i + (size_t)L->id * lut_stride;
                    int q = lower_bound_f(dd, ORC_LUT_LENGTH, wf);
                    double idd, sms;
                    if (q == ORC_LUT_LENGTH) { idd = ii[q - 1]; sms = ss[q - 1]; }
                    else if (q == 0) { idd = ii[0]; sms = ss[0]; }
                    else {
                        double f = (w - dd[q - 1]) / ((double)dd[q] - dd[q - 1]);
                        idd = ii[q - 1] + ((double)ii[q] - ii[q - 1]) * f;
                        sms = ss[q - 1] + ((double)ss[q] - ss[q - 1]) * f;
                    }
                    const float *c = dvp + 5 * (size_t)L->id + 2;
                    double da = dist - w + 0.7 * r80;
                    double sig_air = ((double)c[0] * da + c[1]) * da + c[2];
                    double sig_t = sig_air + sms;
                    double wc = fmin(fmax(w, 0.1), r80 - 0.1);
                    double sh = 2.85 + 0.0014 * r80 * log(wc + 3.0) + 0.06 * wc
                              - 7.4e-5 * wc * wc
                              - (0.22 * r80) / ((wc - r80 - 5.0) * (wc - r80 - 5.0));
                    double hw = 0.052 * log(1.13 + wc / (11.2 - 0.023 * r80))
                              + 0.35 * (0.0017 * r80 * r80 - r80) / ((r80 + 3.0) * (r80 + 3.0) - wc * wc)
                              - 1.61e-9 * wc * (r80 + 3.0) * (r80 + 3.0);
                    double sig_h = sh > 0.0 ? sh : 0.0;
                    double om = fmin(fmax(hw, 0.0), 0.9);
                    double sig_ht = hypot(sig_t, sig_h);
                    double P = (1.0 - om) * idd / (2.0 * ORC_PI * sig_t * sig_t);
                    double H = om * idd / (2.0 * ORC_PI * sig_ht * sig_ht);
                    double ap = sig_t != 0.0 ? -0.5 / (sig_t * sig_t) : -INFINITY;
                    double ah = sig_ht != 0.0 ? -0.5 / (sig_ht * sig_ht) : -INFINITY;
                    double acc = 0.0;
                    for (int s = L->start; s < L->start + L->n; s++) {
                        const orc_spot *S = &spots[s];
                        double tx = (double)S->x / vsadx, ty = (double)S->y / vsady, tz = -1.0;
                        double rx = h[0] - S->x, ry = h[1] - S->y, rz = h[2];
                        double tt = tx * tx + ty * ty + tz * tz;
                        double rr = rx * rx + ry * ry + rz * rz;
                        double dp = tx * rx + ty * ry + tz * rz;
                        double d2 = rr - dp * dp / tt;
                        acc += (double)S->mu * (P * exp(ap * d2) + H * exp(ah * d2));
                    }
                    total += acc;
                    pairs += L->n;
                }
                dose_out[v] = total;
            }
    return pairs;
}

/* ------------------------------------------------------------------ */
/* Photon model (IMRTClasses.cuh:23-57); kernel is [6][6] row = cone  */
/* zenith, columns theta, Am, am, Bm, bm, ray_length                  */
/* (IMRTClasses.cu:290-296, read column-major from the F-order CSV).  */
/* ------------------------------------------------------------------ */
typedef struct {
    int n_profile; const float *prof_r, *prof_i, *prof_s;
    int n_spec; const float *mu_e, *pw, *sw;
    double mu_cal, sad, ssd_s, src_p, src_s, d_mlc, ssw, el_att, esw, el_dmax;
    float kernel[36];
} orc_photon_model;

static double head_transmission_d(const orc_photon_model *m, const float *mlc, int n_pairs,
                                  const double h[3], double Z, double sigma)
{   /* IMRTClasses.cu:120-150; mlc rows {x1, x2, y_offset, y_width} */
    double ms = (m->sad - m->d_mlc) / m->sad;
    double dv = (Z - h[2]) / (m->d_mlc - h[2]);
    double inv = 1.0 / (sigma * sqrt(2.0));
    double tr = 0.0;
    for (int p = 0; p < n_pairs; p++) {
        const float *r = mlc + 4 * p;
        double yb = ((double)r[2] - 0.5 * r[3]) * ms;
        double yt = ((double)r[2] + 0.5 * r[3]) * ms;
        double xl = (double)r[0] * ms, xr = (double)r[1] * ms;
        double t1 = (xl - h[0]) * dv + h[0], t2 = (xr - h[0]) * dv + h[0];
        double e1 = (yb - h[1]) * dv + h[1], e2 = (yt - h[1]) * dv + h[1];
        double ex = 0.5 * (erf(t2 * inv) - erf(t1 * inv));
        double ey = 0.5 * (erf(e2 * inv) - erf(e1 * inv));
        tr += ex * ey;
    }
    return tr;
}

/* termaKernel, IMRTClasses.cu:209-250 (+ offAxisFactors 152-171). */
void orc_terma_m(const orc_geom *g, const orc_photon_model *m, const float *mlc, int n_pairs,
                 const float *wet, double *terma_out, double *electron_out, const uint8_t *mask)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    #pragma omp parallel for schedule(dynamic, 2) collapse(2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t v = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
                if (mask && !mask[v]) { terma_out[v] = 0.0; electron_out[v] = 0.0; continue; }
                float pf[3];
                vox_xyz_f(g, i, j, k, pf);
                double p[3] = { pf[0], pf[1], pf[2] }, h[3];
                img_to_head_d(g, p, h);
                double dps = m->sad - h[2], dss = m->ssd_s - h[2];
                double rc = hypot(h[0], h[1]) * dps / m->sad;
                int q = lower_bound_f(m->prof_r, m->n_profile, (float)rc);
                double oaf, soft;
                if (q == m->n_profile) { oaf = m->prof_i[q - 1]; soft = m->prof_s[q - 1]; }
                else if (q == 0) { oaf = m->prof_i[0]; soft = m->prof_s[0]; }
                else {
                    double f = (rc - m->prof_r[q - 1]) / ((double)m->prof_r[q] - m->prof_r[q - 1]);
                    oaf = m->prof_i[q - 1] + ((double)m->prof_i[q] - m->prof_i[q - 1]) * f;
                    soft = m->prof_s[q - 1] + ((double)m->prof_s[q] - m->prof_s[q - 1]) * f;
                }
                double Tp = head_transmission_d(m, mlc, n_pairs, h, m->sad, m->src_p);
                double Ts = head_transmission_d(m, mlc, n_pairs, h, m->ssd_s, m->src_s);
                double W = wet[v];
                double tr = fmin(1.0, Ts / Tp);
                double terma = 0.0;
                for (int e = 0; e < m->n_spec; e++) {
                    double att = exp(-(double)m->mu_e[e] * W * soft);
                    terma += (1.0 - tr) * (m->pw[e] * att * (m->sad / dps) * (m->sad / dps))
                           + tr * m->sw[e] * att * (m->ssd_s / dss);
                }
                double el = (exp(-m->el_att * W) - exp(-m->el_att * m->el_dmax)) /
                            (1.0 - exp(-m->el_att * m->el_dmax));
                el = fmax(0.0, el); /* fmaxf(0, NaN) = 0, as in C */
                terma_out[v] = oaf * ((1.0 - m->ssw) * Tp * terma + m->ssw * Ts * terma);
                electron_out[v] = m->esw * (0.4 + 0.3 * tr) * el * Tp;
            }
}

void orc_terma(const orc_geom *g, const orc_photon_model *m, const float *mlc, int n_pairs,
               const float *wet, double *terma_out, double *electron_out)
{
    orc_terma_m(g, m, mlc, n_pairs, wet, terma_out, electron_out, NULL);
}

/* OpenMP threads of the oracle's parallel loops (the CPU baseline times one
 * thread and all host cores). */
void orc_set_threads(int n);
int orc_get_threads(void);

/* kernelTilt, IMRTClasses.cu:173-203 (FP64). */
static void kernel_tilt_d(const orc_geom *g, const double p[3], double d[3])
{
    double ux = g->src[0] - p[0], uy = g->src[1] - p[1], uz = g->src[2] - p[2];
    double n = 1.0 / sqrt(ux * ux + uy * uy + uz * uz);
    ux *= n; uy *= n; uz *= n;
    double ax = g->src[0], ay = g->src[1], az = g->src[2];
    double na = 1.0 / sqrt(ax * ax + ay * ay + az * az);
    ax *= na; ay *= na; az *= na;
    double c = ax * ux + ay * uy + az * uz;
    double kx = ay * uz - az * uy, ky = az * ux - ax * uz, kz = ax * uy - ay * ux;
    double rx = ky * d[2] - kz * d[1], ry = kz * d[0] - kx * d[2], rz = kx * d[1] - ky * d[0];
    rx += d[0] * c; ry += d[1] * c; rz += d[2] * c;
    double sc = (kx * d[0] + ky * d[1] + kz * d[2]) / (1.0 + c);
    d[0] = rx + kx * sc; d[1] = ry + ky * sc; d[2] = rz + kz * sc;
}

/* Exported single-vector views of the geometry above, so the tests can pin
 * the rotation chain and the tilt by their invariants.  fp32 != 0 selects
 * the FP32 decision-arithmetic variants (CudaClasses.cu:65-126 with the
 * collimator step of IMRTClasses.cu:41-118); otherwise the FP64 ones. */
void orc_img_to_head(const orc_geom *g, int fp32, const double in[3], double out[3])
{
    if (fp32) {
        float a[3] = { (float)in[0], (float)in[1], (float)in[2] }, b[3];
        img_to_head_f(g, a, b);
        for (int c = 0; c < 3; c++) out[c] = b[c];
    } else {
        img_to_head_d(g, in, out);
    }
}

void orc_head_to_img(const orc_geom *g, int fp32, const double in[3], double out[3])
{
    if (fp32) {
        float a[3] = { (float)in[0], (float)in[1], (float)in[2] }, b[3];
        head_to_img_f(g, a, b);
        for (int c = 0; c < 3; c++) out[c] = b[c];
    } else {
        head_to_img_d(g, in, out);
    }
}

/* kernelTilt of an image-frame vector d for the voxel at image point p. */
void orc_kernel_tilt(const orc_geom *g, const double p[3], double d[3])
{
    kernel_tilt_d(g, p, d);
}

/* Image-frame direction of cone (row r, azimuth az) at image point p, as
 * orc_ccc forms it (IMRTClasses.cu:282-306). */
void orc_cone_dir(const orc_geom *g, const float *kernel36, int r, int az, const double p[3],
                  double d[3])
{
    float thf = kernel36[6 * r + 0] * ORC_PI_F / 180.0f;
    double sth = sin((double)thf), cth = cos((double)thf);
    double phi = (double)((float)az * ORC_PI_F / 6.0f);
    double dh[3] = { sth * cos(phi), sth * sin(phi), cth };
    head_to_img_d(g, dh, d);
    kernel_tilt_d(g, p, d);
}

/* cccKernel, IMRTClasses.cu:252-350.  `terma`, `rho`, `electron` are the
 * FP32 volumes the kernel samples.  Returns the number of cone samples. */
int64_t orc_ccc(const orc_geom *g, const orc_photon_model *m, double mu,
                const float *terma, const float *rho, const float *electron,
                double *dose_out, const uint8_t *mask)
{
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const float sp = g->s / 10.0f;
    /* per-row trip counts follow the FP32 recursion ray_length -= sp */
    int64_t samples = 0;
    #pragma omp parallel for schedule(dynamic, 1) collapse(2) reduction(+:samples)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t v = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
                if (mask && !mask[v]) { dose_out[v] = 0.0; continue; }
                float pf[3], tf[3];
                vox_xyz_f(g, i, j, k, pf);
                tex_f(g, pf, tf);
                double T0 = orc_sample(terma, nx, ny, nz, tf[0], tf[1], tf[2], 0.0);
                if (T0 <= 0.01) { dose_out[v] = 0.0; continue; }
                double p[3] = { pf[0], pf[1], pf[2] };
                double dose = 0.0;
                for (int r = 0; r < 6; r++) {
                    float thf = m->kernel[6 * r + 0] * ORC_PI_F / 180.0f;
                    double A = m->kernel[6 * r + 1], a = m->kernel[6 * r + 2];
                    double B = m->kernel[6 * r + 3], b = m->kernel[6 * r + 4];
                    float L0 = m->kernel[6 * r + 5];
                    double sth = sin((double)thf);
                    for (int az = 0; az < 12; az++) {
                        double d[3];
                        orc_cone_dir(g, m->kernel, r, az, p, d);
                        double Rp = 0.0, Rs = 0.0;
                        for (float L = L0; L >= 0.0f; L = L - sp) {
                            double Lmm = (double)(L * 10.0f);
                            double t[3];
                            for (int c = 0; c < 3; c++)
                                t[c] = (p[c] + d[c] * Lmm + g->iso[c]) / g->s + 0.5;
                            double T = orc_sample(terma, nx, ny, nz, t[0], t[1], t[2], 0.0);
                            double D = orc_sample(rho, nx, ny, nz, t[0], t[1], t[2], ORC_AIR_DENSITY);
                            D = (D > ORC_AIR_DENSITY ? D : (double)ORC_AIR_DENSITY) * (double)sp;
                            double e = exp(-a * D);
                            Rp = Rp * e + T * sth * (A / (a * a)) * (1.0 - e);
                            Rs = Rs * (1.0 - b * D) + T * D * sth * (B / b);
                            samples++;
                        }
                        dose += a * Rp + b * Rs;
                    }
                }
                dose += electron[v];
                dose_out[v] = (dose == dose && dose >= 0.0) ? m->mu_cal * dose * mu : 0.0;
            }
    return samples;
}

/* Number of FP32 steps of `for (L = L0; L >= 0; L -= sp)` (IMRTClasses.cu:316-333). */
int orc_cone_steps(float L0, float spacing_mm)
{
    float sp = spacing_mm / 10.0f;
    int n = 0;
    for (float L = L0; L >= 0.0f; L = L - sp) n++;
    return n;
}

#ifdef _OPENMP
#include <omp.h>
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
#else
void orc_set_threads(int n) { (void)n; }
int orc_get_threads(void) { return 1; }
#endif
