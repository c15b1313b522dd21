/*
 * sanitize_main.c -- drives every entry point of the FP64 oracle on small
 * odd-shaped grids with seeded random inputs, for an AddressSanitizer +
 * UndefinedBehaviorSanitizer build (`make -C oracle sanitize`).  TEST
 * INFRASTRUCTURE ONLY: it checks the oracle's own memory and arithmetic
 * hygiene, not its results (tests/test_oracle*.py do that).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nx, ny, nz;
    float s, inv_s, iso[3], src[3];
    float singa, cosga, sinta, costa, sinca, cosca;
} orc_geom;
typedef struct { float x, y, mu; int id; } orc_spot;
typedef struct { int start, n, id; float r80, energy; } orc_layer;
typedef struct {
    int n_profile; const float *prof_r, *prof_i, *prof_s;
    int n_spec; const float *mu_e, *pw, *sw;
    double mu_cal, sad, ssd_s, src_p, src_s, d_mlc, ssw, el_att, esw, el_dmax;
    float kernel[36];
} orc_photon_model;

void orc_geom_init(orc_geom *, int, int, int, float, const float *, const float *, double, double, double, float);
double orc_sample(const float *, int, int, int, double, double, double, double);
void orc_raytrace_m(const orc_geom *, const float *, double, double *, int32_t *, const uint8_t *);
void orc_smooth(const orc_geom *, const float *, double *);
int orc_make_layers(const float *, int, int, const float *, orc_spot *, orc_layer *);
int64_t orc_pencil(const orc_geom *, float, float, const float *, const orc_spot *, const orc_layer *, int,
                   const float *, const float *, const float *, const float *, int, double *);
void orc_terma_m(const orc_geom *, const orc_photon_model *, const float *, int, const float *, double *,
                 double *, const uint8_t *);
int64_t orc_ccc(const orc_geom *, const orc_photon_model *, double, const float *, const float *, const float *,
                double *, const uint8_t *);
int orc_cone_steps(float, float);
void orc_img_to_head(const orc_geom *, int, const double *, double *);
void orc_head_to_img(const orc_geom *, int, const double *, double *);
void orc_kernel_tilt(const orc_geom *, const double *, double *);
void orc_cone_dir(const orc_geom *, const float *, int, int, const double *, double *);

static uint64_t st = 0x9e3779b97f4a7c15ull;
static double rnd(void)
{
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return (double)(st >> 11) / 9007199254740992.0;
}
static float urand(float lo, float hi) { return lo + (hi - lo) * (float)rnd(); }

static float *vol(int n, float lo, float hi)
{
    float *v = malloc(sizeof(float) * (size_t)n);
    for (int i = 0; i < n; i++) v[i] = urand(lo, hi);
    return v;
}

int main(void)
{
    const int shapes[][3] = {{1, 1, 1}, {2, 3, 1}, {5, 7, 9}, {9, 4, 6}, {3, 11, 2}};
    const double angles[][3] = {{0, 0, 0}, {37, 12, 25}, {270, 300, 90}, {180, 45, 355}};
    const int E = 4, LS = 403;   /* energies, LUT row stride (> 400) */
    float *dvp = malloc(sizeof(float) * 5 * E), *ld = malloc(sizeof(float) * E * LS);
    float *lsg = malloc(sizeof(float) * E * LS), *lid = malloc(sizeof(float) * E * LS);
    for (int e = 0; e < E; e++) {
        dvp[5 * e] = 70.0f + 10 * e; dvp[5 * e + 1] = e == 0 ? -0.65f : 20.0f + 30 * e;
        dvp[5 * e + 2] = 1e-6f; dvp[5 * e + 3] = 1e-3f; dvp[5 * e + 4] = 3.0f;
        for (int q = 0; q < LS; q++) {
            ld[e * LS + q] = (float)q; lsg[e * LS + q] = 0.01f * q; lid[e * LS + q] = 1.0f + 0.002f * q;
        }
    }
    float pr[6] = {0, 20, 60, 120, 200, 300}, pi[6] = {1, 1, 0.98f, 0.9f, 0.5f, 0.1f}, ps[6] = {1, 1, 1, 1.1f, 1.2f, 1.3f};
    float mu_e[3] = {0.05f, 0.03f, 0.02f}, pw[3] = {0.3f, 0.4f, 0.3f}, sw[3] = {0.2f, 0.5f, 0.3f};
    orc_photon_model pm = {6, pr, pi, ps, 3, mu_e, pw, sw, 0.008, 1000, 850, 0.6, 20, 500, 0.08, 0.5, 0.01, 15,
                           {0}};
    const float th[6] = {3.75f, 11.25f, 18.75f, 26.25f, 33.75f, 41.25f};
    for (int r = 0; r < 6; r++) {
        pm.kernel[6 * r] = th[r]; pm.kernel[6 * r + 1] = 0.01f; pm.kernel[6 * r + 2] = 1.5f + 0.1f * r;
        pm.kernel[6 * r + 3] = 0.002f; pm.kernel[6 * r + 4] = 0.3f; pm.kernel[6 * r + 5] = 1.2f + 0.4f * r;
    }
    long checks = 0;
    for (size_t a = 0; a < sizeof(shapes) / sizeof(shapes[0]); a++) {
        const int nz = shapes[a][0], ny = shapes[a][1], nx = shapes[a][2], n = nx * ny * nz;
        for (size_t b = 0; b < sizeof(angles) / sizeof(angles[0]); b++) {
            const float s = urand(1.5f, 3.0f);
            const float org[3] = {urand(-5, 5), urand(-5, 5), urand(-5, 5)};
            const float iso[3] = {org[0] + 0.5f * s * nx, org[1] + 0.5f * s * ny, org[2] + 0.5f * s * nz};
            orc_geom g;
            orc_geom_init(&g, nz, ny, nx, s, org, iso, angles[b][0], angles[b][1], angles[b][2], 1000.0f);
            float *rho = vol(n, 0.0f, 1.8f);
            uint8_t *mask = malloc((size_t)n);
            for (int v = 0; v < n; v++) mask[v] = (uint8_t)(rnd() < 0.6);
            double *wet = malloc(sizeof(double) * n), *out = malloc(sizeof(double) * n);
            double *el = malloc(sizeof(double) * n);
            int32_t *ns = malloc(sizeof(int32_t) * n);
            orc_raytrace_m(&g, rho, 0.0012, wet, ns, NULL);
            orc_raytrace_m(&g, rho, 0.0, wet, ns, mask);
            float *w32 = malloc(sizeof(float) * n);
            for (int v = 0; v < n; v++) w32[v] = (float)wet[v];
            orc_smooth(&g, w32, out);
            /* spots: ids in and out of range, negative, unsorted */
            const int nsp = 17;
            float *rows = malloc(sizeof(float) * 4 * nsp);
            for (int q = 0; q < nsp; q++) {
                rows[4 * q] = urand(-10, 10); rows[4 * q + 1] = urand(-10, 10); rows[4 * q + 2] = urand(0.5f, 2.0f);
                rows[4 * q + 3] = (float)((int)(rnd() * (E + 3)) - 1);
            }
            orc_spot spots[17];
            orc_layer layers[4];
            int nl = orc_make_layers(rows, nsp, E, dvp, spots, layers);
            for (int v = 0; v < n; v++) w32[v] = urand(-0.05f, 9.0f);
            orc_pencil(&g, 1340.0f, 1930.0f, w32, spots, layers, nl, dvp, ld, lsg, lid, LS, out);
            /* MLC: closed pairs, equal tips, inverted pairs */
            const int np = 13;
            float *mlc = malloc(sizeof(float) * 4 * np);
            for (int p = 0; p < np; p++) {
                float x1 = urand(-40, 20), x2 = x1 + urand(-5, 40);
                if (p % 4 == 0) x2 = x1;
                mlc[4 * p] = x1; mlc[4 * p + 1] = x2; mlc[4 * p + 2] = -30.0f + 5.0f * p; mlc[4 * p + 3] = 5.0f;
            }
            orc_terma_m(&g, &pm, mlc, np, w32, wet, el, NULL);
            orc_terma_m(&g, &pm, mlc, np, w32, wet, el, mask);
            float *t32 = malloc(sizeof(float) * n), *e32 = malloc(sizeof(float) * n);
            for (int v = 0; v < n; v++) { t32[v] = (float)wet[v] + urand(0, 0.05f); e32[v] = (float)el[v]; }
            orc_ccc(&g, &pm, 100.0, t32, rho, e32, out, NULL);
            orc_ccc(&g, &pm, 50.0, t32, rho, e32, out, mask);
            double p[3] = {urand(-20, 20), urand(-20, 20), urand(-20, 20)}, h[3], q[3], d[3] = {0.3, -0.2, 0.9};
            orc_img_to_head(&g, 0, p, h); orc_head_to_img(&g, 1, h, q);
            orc_kernel_tilt(&g, p, d);
            for (int r = 0; r < 6; r++)
                for (int az = 0; az < 12; az++) orc_cone_dir(&g, pm.kernel, r, az, p, d);
            for (int t = 0; t < 64; t++)
                (void)orc_sample(rho, nx, ny, nz, urand(-3, nx + 3), urand(-3, ny + 3), urand(-3, nz + 3), 0.5);
            checks++;
            free(rho); free(mask); free(wet); free(out); free(el); free(ns); free(w32); free(rows); free(mlc);
            free(t32); free(e32);
        }
    }
    for (int k = 0; k < 6; k++) (void)orc_cone_steps(pm.kernel[6 * k + 5], 2.5f);
    free(dvp); free(ld); free(lsg); free(lid);
    printf("sanitize: ok (%ld grid/angle cases)\n", checks);
    return 0;
}
