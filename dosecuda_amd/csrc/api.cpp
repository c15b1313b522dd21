// api.cpp -- the C ABI of libdosehip.so (include/dosehip.h).
//
// Host orchestration of the reference (DoseCUDA/dose_kernels/IMPTClasses.cu
// :274-340, IMRTClasses.cu:353-405) and the numeric part of its Python
// boundary (dosemodule.cu:104-557): angle and isocentre adjustment, spot sort
// and layer import, MLC edge scaling, beam-model tables.  Differences by
// design: scratch buffers persist per device (no malloc/free per call), the
// density is uploaded once per call into a padded layout instead of a texture
// array, and every kernel launch is checked.
#include "dh_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

// ---------------------------------------------------------------------------
// error plumbing
// ---------------------------------------------------------------------------
static thread_local std::string t_err;

int dh_set_error(int code, const char *msg)
{
    t_err = msg ? msg : "";
    return code;
}

#define guarded dh_guarded

extern "C" const char *dh_last_error(void) { return t_err.c_str(); }
extern "C" int dh_version(void) { return 10000; }

extern "C" int dh_device_count(int *count)
{
    return guarded([&] {
        int n = 0;
        DH_HIP(hipGetDeviceCount(&n));
        if (count) *count = n;
    });
}

void dh_noalloc_setup(int dev);
extern "C" int dh_init(int gpu_id)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        dh_noalloc_setup(gpu_id);   // device address of the fixed scratch array
    });
}

void dh_device_release(int dev);
extern "C" int dh_release(int gpu_id)
{
    return guarded([&] { dh_device_release(gpu_id); });
}

static size_t nvox_of(const dh_grid *g)
{
    DH_REQUIRE(g, "grid must not be NULL");
    DH_REQUIRE(g->nx > 0 && g->ny > 0 && g->nz > 0, "volume dimensions must be positive");
    DH_REQUIRE(std::isfinite(g->spacing) && g->spacing > 0.0f, "spacing[0] must be positive");
    size_t n = (size_t)g->nx * g->ny * g->nz;
    size_t np = (size_t)(g->nx + 3) * (g->ny + 3) * (g->nz + 3);
    DH_REQUIRE(np < (size_t)INT32_MAX, "volume too large for 32-bit voxel indices");
    return n;
}

static float proton_source_distance(const dh_proton_beam *b)
{   // IMPTClasses.cuh:36, with the doubles stored as floats (dosemodule.cu:168-169)
    return ((float)b->vsadx + (float)b->vsady) / 2.0f;
}

// ---------------------------------------------------------------------------
// proton: WET
// ---------------------------------------------------------------------------
dh_geom dh_proton_geom(const dh_grid *grid, const dh_proton_beam *beam)
{
    nvox_of(grid);
    DH_REQUIRE(beam, "null beam");
    dh_geom g = dh_make_geom(*grid, beam->iso, beam->gantry_angle, beam->couch_angle, 0.0,
                             proton_source_distance(beam));
    g.sinca = 0.0f; g.cosca = 1.0f;   // protons have no collimator rotation
    return g;
}

static void proton_wet_device(int dev, const dh_grid *grid, const dh_proton_beam *beam,
                              const float *d_rlsp, float *d_wet_out, float *d_wet_raw,
                              hipStream_t s)
{
    size_t n = nvox_of(grid);
    DH_REQUIRE(beam && d_rlsp && d_wet_out, "null argument");
    dh_geom g = dh_proton_geom(grid, beam);
    dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
    float4 *vpad = (float4 *)dh_scratch(dev, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
    float *raw = d_wet_raw ? d_wet_raw : (float *)dh_scratch(dev, DH_SLOT_WET_RAW, sizeof(float) * n);
    dh_launch_pad_quad(d_rlsp, vpad, g.nx, g.ny, g.nz, 0.0f, s);      // IMPTClasses.cu:292, border 0
    dh_launch_raytrace(g, vpad, raw, s);
    dh_launch_smooth(g, raw, d_wet_out, s);
}

extern "C" int dh_proton_raytrace_dev(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                      const float *d_rlsp, float *d_wet_out, float *d_wet_raw_out,
                                      void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        proton_wet_device(gpu_id, grid, beam, d_rlsp, d_wet_out, d_wet_raw_out, (hipStream_t)stream);
    });
}

extern "C" int dh_proton_raytrace(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                  const float *rlsp, float *wet_out)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        size_t n = nvox_of(grid);
        DH_REQUIRE(rlsp && wet_out, "null buffer");
        hipStream_t s = dh_stream(gpu_id);
        float *d_in = (float *)dh_scratch(gpu_id, DH_SLOT_IN0, sizeof(float) * n);
        float *d_out = (float *)dh_scratch(gpu_id, DH_SLOT_OUT0, sizeof(float) * n);
        DH_HIP(hipMemcpyAsync(d_in, rlsp, sizeof(float) * n, hipMemcpyHostToDevice, s));
        proton_wet_device(gpu_id, grid, beam, d_in, d_out, nullptr, s);
        DH_HIP(hipMemcpyAsync(wet_out, d_out, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        DH_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int dh_stage_smooth_dev(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                   const float *d_wet_raw, float *d_wet_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        nvox_of(grid);
        DH_REQUIRE(beam && d_wet_raw && d_wet_out, "null argument");
        dh_geom g = dh_make_geom(*grid, beam->iso, beam->gantry_angle, beam->couch_angle, 0.0,
                                 proton_source_distance(beam));
        g.sinca = 0.0f; g.cosca = 1.0f;
        dh_launch_smooth(g, d_wet_raw, d_wet_out, (hipStream_t)stream);
    });
}

// HU -> RLSP / density on the device (plan_impt.py:112-122, plan_imrt.py:114-124).
extern "C" int dh_hu_convert_dev(int gpu_id, const void *d_hu, int32_t hu_is_double, int64_t n,
                                 const double *xp, const double *fp, int32_t m, float *d_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        DH_REQUIRE(n >= 0, "negative voxel count");
        DH_REQUIRE(n == 0 || (d_hu && d_out), "null argument");
        DH_REQUIRE(xp && fp, "null HU table");
        DH_REQUIRE(m >= 1 && m <= DH_HU_TABLE_MAX, "HU table must have 1..128 nodes");
        dh_hu_table t;
        t.m = m;
        for (int i = 0; i < m; i++) {
            t.x[i] = xp[i];
            t.y[i] = fp[i];
        }
        dh_launch_hu_table(t, d_hu, hu_is_double != 0, n, d_out, (hipStream_t)stream);
    });
}

extern "C" int dh_stage_raytrace_dev(int gpu_id, const dh_grid *grid, const float iso[3],
                                     double gantry, double couch, float src_dist, float border,
                                     const float *d_vol, float *d_wet_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        nvox_of(grid);
        DH_REQUIRE(iso && d_vol && d_wet_out, "null argument");
        dh_geom g = dh_make_geom(*grid, iso, gantry, couch, 0.0, src_dist);
        dh_pad p = dh_make_pad(g.nx, g.ny, g.nz);
        float4 *vpad = (float4 *)dh_scratch(gpu_id, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
        dh_launch_pad_quad(d_vol, vpad, g.nx, g.ny, g.nz, border, (hipStream_t)stream);
        dh_launch_raytrace(g, vpad, d_wet_out, (hipStream_t)stream);
    });
}

// ---------------------------------------------------------------------------
// proton: spots, layers and the pencil-beam dose
// ---------------------------------------------------------------------------
// make_spot_array + importLayers (dosemodule.cu:211-225, IMPTClasses.cu:29-67).
// The sort is stable, so spots of one layer keep the caller's order (the
// reference's std::sort leaves that order unspecified).
dh_proton_tables dh_build_proton_tables(const dh_proton_beam *beam, const dh_proton_lut *lut,
                                        const float *spot_rows, int32_t n_spots)
{
    DH_REQUIRE(beam && lut, "null argument");
    DH_REQUIRE(n_spots >= 0, "negative spot count");
    DH_REQUIRE(n_spots == 0 || spot_rows, "null spot array");
    DH_REQUIRE(lut->n_energies >= 0, "negative energy count");
    DH_REQUIRE(lut->n_energies == 0 ||
                   (lut->divergence_params && lut->depths && lut->sigmas && lut->idds),
               "null beam-model table");
    DH_REQUIRE(lut->row_stride >= DH_LUT_LENGTH || lut->n_energies == 0,
               "lookup tables need at least 400 columns");
    struct sp { float x, y, mu; int id, row; };
    std::vector<sp> s((size_t)n_spots);
    for (int32_t q = 0; q < n_spots; q++) {
        s[q].x = spot_rows[4 * q];
        s[q].y = spot_rows[4 * q + 1];
        s[q].mu = spot_rows[4 * q + 2];
        s[q].id = static_cast<int>(spot_rows[4 * q + 3]);
        s[q].row = q;
    }
    std::stable_sort(s.begin(), s.end(), [](const sp &a, const sp &b) { return a.id < b.id; });

    dh_proton_tables t;
    const float vx = (float)beam->vsadx, vy = (float)beam->vsady;
    t.spots.resize((size_t)8 * s.size());
    t.order.resize(s.size());
    for (size_t q = 0; q < s.size(); q++) {
        t.order[q] = s[q].row;
        // unit spot-axis direction (x, y components), IMPTClasses.cu:125
        float tx = s[q].x / vx, ty = s[q].y / vy;
        float n = (float)(1.0 / std::sqrt((double)tx * tx + (double)ty * ty + 1.0));
        float *r = &t.spots[8 * q];
        r[0] = s[q].x; r[1] = s[q].y; r[2] = tx * n; r[3] = ty * n;
        r[4] = s[q].mu; r[5] = 0.0f; r[6] = 0.0f; r[7] = 0.0f;
    }
    int start = 0;
    for (int e = 0; e < lut->n_energies; e++) {
        int end = start;
        while (end < n_spots && s[end].id == e) end++;
        int cnt = end - start;
        if (!cnt) continue;
        const float *dvp = lut->divergence_params + (size_t)5 * e;   // dvp_len = 5
        dh_layer_k L;
        L.start = start; L.n = cnt;
        L.r80 = dvp[1];
        L.cut = 1.1f * dvp[1];                                        // IMPTClasses.cu:224
        L.A = dvp[2]; L.B = dvp[3]; L.C = dvp[4]; L.pad_ = 0.0f;
        t.layers.push_back(L);
        const size_t off = (size_t)e * lut->row_stride;
        for (int q = 0; q < DH_LUT_LENGTH; q++) {
            float idd = lut->idds[off + q];
            if (!(idd >= 0.0f) || !std::isfinite(idd)) t.fold = false;
            t.luts.push_back(make_float4(lut->depths[off + q], lut->sigmas[off + q], idd, 0.0f));
        }
        start += cnt;
    }
    t.n_layers = (int)t.layers.size();
    return t;
}

// Device copy of the proton tables; stays alive for the plan's lifetime.
struct dh_proton_plan {
    int dev = 0;
    dh_grid grid;
    dh_proton_beam beam;
    dh_geom g;
    int n_layers = 0;
    bool fold = true;
    dh_layer_k *d_layers = nullptr;
    float4 *d_luts = nullptr;
    float *d_spots = nullptr;
    std::vector<dh_layer_k> layers;
};

static void free_plan(dh_proton_plan *p)
{
    if (!p) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(p->dev);
    if (p->d_layers) (void)hipFree(p->d_layers);
    if (p->d_luts) (void)hipFree(p->d_luts);
    if (p->d_spots) (void)hipFree(p->d_spots);
    (void)hipSetDevice(cur);
    delete p;
}

static dh_proton_plan *make_plan(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                 const dh_proton_lut *lut, const float *spot_rows, int32_t n_spots)
{
    nvox_of(grid);
    dh_device_enter(gpu_id);
    dh_proton_tables t = dh_build_proton_tables(beam, lut, spot_rows, n_spots);
    dh_proton_plan *p = new dh_proton_plan();
    p->dev = gpu_id;
    p->grid = *grid;
    p->beam = *beam;
    p->g = dh_make_geom(*grid, beam->iso, beam->gantry_angle, beam->couch_angle, 0.0,
                        proton_source_distance(beam));
    p->g.sinca = 0.0f; p->g.cosca = 1.0f;
    p->n_layers = t.n_layers;
    p->fold = t.fold;
    p->layers = t.layers;
    try {
        if (t.n_layers) {
            DH_HIP(hipMalloc(&p->d_layers, sizeof(dh_layer_k) * t.layers.size()));
            DH_HIP(hipMalloc(&p->d_luts, sizeof(float4) * t.luts.size()));
            DH_HIP(hipMalloc(&p->d_spots, sizeof(float) * t.spots.size()));
            DH_HIP(hipMemcpy(p->d_layers, t.layers.data(), sizeof(dh_layer_k) * t.layers.size(), hipMemcpyHostToDevice));
            DH_HIP(hipMemcpy(p->d_luts, t.luts.data(), sizeof(float4) * t.luts.size(), hipMemcpyHostToDevice));
            DH_HIP(hipMemcpy(p->d_spots, t.spots.data(), sizeof(float) * t.spots.size(), hipMemcpyHostToDevice));
        }
    } catch (...) {
        free_plan(p);
        throw;
    }
    return p;
}

static void plan_dose(dh_proton_plan *p, const float *d_wet, float *d_dose, hipStream_t s)
{
    dh_launch_pencil(p->g, (float)p->beam.vsadx, (float)p->beam.vsady, p->n_layers, p->d_layers,
                     p->d_luts, p->d_spots, p->fold, d_wet, d_dose, s);
}

extern "C" int dh_proton_plan_create(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                     const dh_proton_lut *lut, const float *spot_rows,
                                     int32_t n_spots, dh_proton_plan **out)
{
    return guarded([&] {
        DH_REQUIRE(out, "null output handle");
        *out = make_plan(gpu_id, grid, beam, lut, spot_rows, n_spots);
    });
}

extern "C" int dh_proton_plan_layers(const dh_proton_plan *p, int32_t *n_layers,
                                     int32_t *layer_spots, float *layer_cut)
{
    return guarded([&] {
        DH_REQUIRE(p, "null plan");
        if (n_layers) *n_layers = p->n_layers;
        for (int l = 0; l < p->n_layers; l++) {
            if (layer_spots) layer_spots[l] = p->layers[l].n;
            if (layer_cut) layer_cut[l] = p->layers[l].cut;
        }
    });
}

extern "C" int dh_proton_plan_dose_dev(dh_proton_plan *p, const float *d_wet, float *d_dose,
                                       void *stream)
{
    return guarded([&] {
        DH_REQUIRE(p && d_wet && d_dose, "null argument");
        dh_device_enter(p->dev);
        plan_dose(p, d_wet, d_dose, (hipStream_t)stream);
    });
}

extern "C" int dh_proton_plan_destroy(dh_proton_plan *p)
{
    return guarded([&] { free_plan(p); });
}

extern "C" int dh_proton_spot(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                              const dh_proton_lut *lut, const float *spot_rows, int32_t n_spots,
                              const float *wet, float *dose_out)
{
    dh_proton_plan *p = nullptr;
    int rc = guarded([&] {
        size_t n = nvox_of(grid);
        DH_REQUIRE(wet && dose_out, "null buffer");
        p = make_plan(gpu_id, grid, beam, lut, spot_rows, n_spots);
        hipStream_t s = dh_stream(gpu_id);
        float *d_in = (float *)dh_scratch(gpu_id, DH_SLOT_IN0, sizeof(float) * n);
        float *d_out = (float *)dh_scratch(gpu_id, DH_SLOT_OUT0, sizeof(float) * n);
        DH_HIP(hipMemcpyAsync(d_in, wet, sizeof(float) * n, hipMemcpyHostToDevice, s));
        plan_dose(p, d_in, d_out, s);
        DH_HIP(hipMemcpyAsync(dose_out, d_out, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        DH_HIP(hipStreamSynchronize(s));
    });
    if (p) {
        std::string keep = t_err;
        free_plan(p);
        t_err = keep;
    }
    return rc;
}

// ---------------------------------------------------------------------------
// photon: TERMA + CCCS
// ---------------------------------------------------------------------------
static dh_geom photon_geom(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m)
{
    return dh_make_geom(*grid, cp->iso, cp->ga, cp->ta, cp->ca, (float)m->primary_source_distance);
}

static void check_model(const dh_photon_model *m)
{
    DH_REQUIRE(m, "null photon model");
    DH_REQUIRE(m->n_profile > 0 && m->n_profile <= DH_MAX_PROFILE, "profile must have 1..32 points");
    DH_REQUIRE(m->n_spectral >= 0 && m->n_spectral <= DH_MAX_SPECTRAL, "spectrum must have at most 16 energies");
    DH_REQUIRE(m->profile_radius && m->profile_intensities && m->profile_softening, "null profile");
    DH_REQUIRE(m->n_spectral == 0 || (m->spectrum_attenuation_coefficients &&
               m->spectrum_primary_weights && m->spectrum_scatter_weights), "null spectrum");
}

static dh_terma_args make_terma_args(const dh_grid *grid, const dh_photon_cp *cp,
                                     const dh_photon_model *m, const float *mlc_rows, int32_t n_pairs)
{
    check_model(m);
    DH_REQUIRE(n_pairs >= 0 && n_pairs <= DH_MAX_PAIRS, "at most 128 MLC leaf pairs are supported");
    DH_REQUIRE(n_pairs == 0 || mlc_rows, "null MLC array");
    dh_terma_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = photon_geom(grid, cp, m);
    const float sad = (float)m->primary_source_distance, dm = (float)m->mlc_distance;
    const float ms = (sad - dm) / sad;                       // IMRTClasses.cu:122
    int q = 0;
    for (int i = 0; i < n_pairs; i++) {
        const float *r = mlc_rows + 4 * i;
        // closed pairs (x1 == x2) contribute exactly 0 * ey in FP32
        if (r[0] == r[1]) continue;
        a.mlc[q++] = make_float4(r[0] * ms, r[1] * ms, (r[2] - 0.5f * r[3]) * ms,
                                 (r[2] + 0.5f * r[3]) * ms);
    }
    // Consecutive open pairs with equal leaf tips whose y ranges meet edge to
    // edge become one rectangle: the head transmission sums such a run as
    // ex * (erf(y_top) - erf(y_bottom)) / 2 (the per-pair terms telescope),
    // which needs only the run's outer edges.  A rectangular field is one.
    int nr = 0;
    for (int i = 0; i < q; i++) {
        if (nr > 0 && a.mlc[i].x == a.mlc[nr - 1].x && a.mlc[i].y == a.mlc[nr - 1].y &&
            a.mlc[i].z == a.mlc[nr - 1].w)
            a.mlc[nr - 1].w = a.mlc[i].w;
        else
            a.mlc[nr++] = a.mlc[i];
    }
    q = nr;
    a.n_pairs = q;
    a.sorted = 1;
    for (int i = 1; i < q; i++)
        if (!(a.mlc[i].z >= a.mlc[i - 1].z && a.mlc[i].w >= a.mlc[i - 1].w)) a.sorted = 0;
    a.sad = sad;
    a.ssd_s = (float)m->scatter_source_distance;
    a.inv_p = 1.0f / ((float)m->primary_source_size * sqrtf(2.f));     // IMRTClasses.cu:123-124
    a.inv_s_src = 1.0f / ((float)m->scatter_source_size * sqrtf(2.f));
    a.d_mlc = dm;
    a.ssw = (float)m->scatter_source_weight;
    a.el_att = (float)m->electron_attenuation;
    a.esw = (float)m->electron_source_weight;
    a.el_norm_a = expf(-a.el_att * (float)m->electron_fitted_dmax);
    a.el_norm_b = 1.0f - a.el_norm_a;
    a.n_prof = m->n_profile;
    for (int i = 0; i < m->n_profile; i++) {
        a.prof_r[i] = m->profile_radius[i];
        a.prof_i[i] = m->profile_intensities[i];
        a.prof_s[i] = m->profile_softening[i];
    }
    a.n_spec = m->n_spectral;
    for (int i = 0; i < m->n_spectral; i++) {
        a.mu_e[i] = m->spectrum_attenuation_coefficients[i];
        a.pw[i] = m->spectrum_primary_weights[i];
        a.sw[i] = m->spectrum_scatter_weights[i];
    }
    return a;
}

static dh_ccc_args make_ccc_args(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m)
{
    check_model(m);
    dh_ccc_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = photon_geom(grid, cp, m);
    a.p = dh_make_pad(a.g.nx, a.g.ny, a.g.nz);
    dh_uvec(a.g, 0.0f, 0.0f, 0.0f, a.ax, a.ay, a.az);          // IMRTClasses.cu:177-179
    a.sp = a.g.s / 10.0f;                                      // IMRTClasses.cu:278
    a.mu_cal = (float)m->mu_calibration;
    a.mu = (float)cp->mu;
    for (int r = 0; r < 6; r++) {
        const float *K = m->kernel + 6 * r;
        const float th = K[0] * DH_PI_F / 180.0f;
        const float sth = sinf(th), cth = cosf(th);
        const float A = K[1], am = K[2], B = K[3], bm = K[4];
        a.L0[r] = K[5];
        a.a[r] = am;
        a.b[r] = bm;
        a.c1[r] = sth * (A / (am * am));
        a.c2p[r] = sth * (B / bm) * a.sp;
        a.bp[r] = bm * a.sp;
        a.a2[r] = -am * a.sp * DH_LOG2E_F;
        int ns = 0;
        for (float L = K[5]; L >= 0.0f; L = L - a.sp) {
            ns++;
            DH_REQUIRE(ns < 100000, "cone ray_length / spacing gives too many steps");
        }
        a.nsteps[r] = ns;
        for (int az = 0; az < 12; az++) {
            const float ang = (float)az * DH_PI_F / 6.0f;      // IMRTClasses.cu:286-288
            const float sx = sinf(ang), cx = cosf(ang);
            float dx, dy, dz;
            dh_head_to_img(a.g, sth * cx, sth * sx, cth, dx, dy, dz);
            a.dir[3 * (r * 12 + az) + 0] = dx;
            a.dir[3 * (r * 12 + az) + 1] = dy;
            a.dir[3 * (r * 12 + az) + 2] = dz;
        }
    }
    return a;
}

// Unit vector source -> isocentre of a control point in volume axes (the
// beam direction channels of the surrogate input).
static void sur_direction(const dh_photon_cp *cp, float dir[3])
{
    const double ga = std::fmod(cp->ga + 180.0, 360.0) * (M_PI / 180.0), ta = cp->ta * (M_PI / 180.0);
    const double xg = -std::sin(ga), yg = std::cos(ga);
    const double sx = xg * std::cos(ta), sy = yg, sz = -xg * std::sin(ta);
    const double nrm = std::sqrt(sx * sx + sy * sy + sz * sz);
    dir[0] = (float)(-sx / nrm);
    dir[1] = (float)(-sy / nrm);
    dir[2] = (float)(-sz / nrm);
}

// Which padded volume and radiological depth each (device, workspace) holds
// from its last photon ray trace, so a batch can reuse them for the next
// control point of the same case with the same source position.
struct wet_memo {
    const void *vpad = nullptr, *wet = nullptr;
};
static wet_memo g_wet_memo[64][DH_MAX_WORKSPACES];

static void photon_device(int dev, const dh_grid *grid, const dh_photon_cp *cp,
                          const dh_photon_model *m, const float *mlc_rows, int32_t n_pairs,
                          const float *d_density, float *d_dose, float *d_terma,
                          float *d_electron, float *d_wet, hipStream_t s,
                          const dh_sur_net *net = nullptr, int stride = 3, float tol = 0.01f, int dilate = 2,
                          bool reuse_wet = false)
{
    size_t n = nvox_of(grid);
    DH_REQUIRE(cp && d_density && d_dose, "null argument");
    dh_terma_args ta = make_terma_args(grid, cp, m, mlc_rows, n_pairs);
    dh_ccc_args ca = make_ccc_args(grid, cp, m);
    dh_pad p = ca.p;
    float4 *vpad = (float4 *)dh_scratch(dev, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
    float *wet = d_wet ? d_wet : (float *)dh_scratch(dev, DH_SLOT_WET_RAW, sizeof(float) * n);
    float *terma = d_terma ? d_terma : (float *)dh_scratch(dev, DH_SLOT_TMP0, sizeof(float) * n);
    float *el = d_electron ? d_electron : (float *)dh_scratch(dev, DH_SLOT_TMP1, sizeof(float) * n);
    void *trp = dh_scratch(dev, DH_SLOT_TR_PAD, dh_tr_pad_bytes(ca));
    int *act = (int *)dh_scratch(dev, DH_SLOT_ACTIVE, sizeof(int) * n);
    int *cnt = (int *)dh_scratch(dev, DH_SLOT_COUNTER, 256);
    wet_memo &memo = g_wet_memo[dev & 63][dh_get_workspace()];
    // The radiological depth depends on the density, the grid and the source
    // position only (not on the collimator or the MLC): a caller that knows
    // the previous control point on this workspace had the same ones may ask
    // to reuse it; it is reused only if this workspace still holds it.
    if (!(reuse_wet && !d_wet && memo.vpad == vpad && memo.wet == wet)) {
        dh_launch_pad_quad(d_density, vpad, ca.g.nx, ca.g.ny, ca.g.nz, DH_AIR_DENSITY, s);  // IMRTClasses.cu:394
        dh_launch_raytrace(ca.g, vpad, wet, s);
    }
    memo.vpad = d_wet ? nullptr : vpad;
    memo.wet = d_wet ? nullptr : wet;
    dh_launch_terma(ta, wet, terma, el, s);
    if (net) {
        float dir[3];
        sur_direction(cp, dir);
        dh_launch_sur_dose(net, ca, terma, d_density, el, dir, stride, tol, dilate, d_dose, s);
    } else {
        dh_launch_ccc(ca, terma, d_density, el, trp, act, cnt, d_dose, s);
    }
}

extern "C" int dh_photon_dose_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                  const dh_photon_model *model, const float *mlc_rows,
                                  int32_t n_pairs, const float *d_density, float *d_dose_out,
                                  float *d_terma_out, float *d_electron_out, float *d_wet_out,
                                  void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        photon_device(gpu_id, grid, cp, model, mlc_rows, n_pairs, d_density, d_dose_out,
                      d_terma_out, d_electron_out, d_wet_out, (hipStream_t)stream);
    });
}

extern "C" int dh_photon_dose(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                              const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                              const float *density, float *dose_out, float *terma_out,
                              float *electron_out)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        size_t n = nvox_of(grid);
        DH_REQUIRE(density && dose_out, "null buffer");
        hipStream_t s = dh_stream(gpu_id);
        float *d_in = (float *)dh_scratch(gpu_id, DH_SLOT_IN0, sizeof(float) * n);
        float *d_dose = (float *)dh_scratch(gpu_id, DH_SLOT_OUT0, sizeof(float) * n);
        float *d_t = (float *)dh_scratch(gpu_id, DH_SLOT_OUT1, sizeof(float) * n);
        float *d_e = (float *)dh_scratch(gpu_id, DH_SLOT_OUT2, sizeof(float) * n);
        DH_HIP(hipMemcpyAsync(d_in, density, sizeof(float) * n, hipMemcpyHostToDevice, s));
        photon_device(gpu_id, grid, cp, model, mlc_rows, n_pairs, d_in, d_dose, d_t, d_e, nullptr, s);
        DH_HIP(hipMemcpyAsync(dose_out, d_dose, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        if (terma_out) DH_HIP(hipMemcpyAsync(terma_out, d_t, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        if (electron_out) DH_HIP(hipMemcpyAsync(electron_out, d_e, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        DH_HIP(hipStreamSynchronize(s));
    });
}

extern "C" int dh_stage_terma_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                  const dh_photon_model *model, const float *mlc_rows,
                                  int32_t n_pairs, const float *d_wet, float *d_terma_out,
                                  float *d_electron_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        nvox_of(grid);
        DH_REQUIRE(cp && d_wet && d_terma_out && d_electron_out, "null argument");
        dh_terma_args ta = make_terma_args(grid, cp, model, mlc_rows, n_pairs);
        dh_launch_terma(ta, d_wet, d_terma_out, d_electron_out, (hipStream_t)stream);
    });
}

extern "C" int dh_stage_ccc_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                const dh_photon_model *model, const float *d_terma,
                                const float *d_density, const float *d_electron,
                                float *d_dose_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        size_t n = nvox_of(grid);
        DH_REQUIRE(cp && d_terma && d_density && d_electron && d_dose_out, "null argument");
        dh_ccc_args ca = make_ccc_args(grid, cp, model);
        dh_pad p = ca.p;
        void *trp = dh_scratch(gpu_id, DH_SLOT_TR_PAD, dh_tr_pad_bytes(ca));
        int *act = (int *)dh_scratch(gpu_id, DH_SLOT_ACTIVE, sizeof(int) * n);
        int *cnt = (int *)dh_scratch(gpu_id, DH_SLOT_COUNTER, 256);
        dh_launch_ccc(ca, d_terma, d_density, d_electron, trp, act, cnt, d_dose_out, (hipStream_t)stream);
    });
}

extern "C" int dh_stage_ccc_list_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                     const dh_photon_model *model, const float *d_terma,
                                     const float *d_density, const float *d_electron,
                                     const int32_t *d_list, int32_t n, float *d_dose_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        size_t nv = nvox_of(grid);
        DH_REQUIRE(cp && d_terma && d_density && d_electron && d_dose_out, "null argument");
        DH_REQUIRE(n >= 0 && (size_t)n <= nv && (n == 0 || d_list), "bad voxel list");
        dh_ccc_args ca = make_ccc_args(grid, cp, model);
        dh_pad p = ca.p;
        void *trp = dh_scratch(gpu_id, DH_SLOT_TR_PAD, dh_tr_pad_bytes(ca));
        int *cnt = (int *)dh_scratch(gpu_id, DH_SLOT_COUNTER, 256);
        dh_launch_ccc_list(ca, d_terma, d_density, d_electron, trp, d_list, n, cnt, d_dose_out,
                           (hipStream_t)stream);
    });
}

extern "C" int dh_sur_check_work_size(const dh_grid *grid, int32_t stride, int64_t *n_ints)
{
    return guarded([&] {
        DH_REQUIRE(grid && n_ints && stride >= 1, "bad argument");
        nvox_of(grid);
        *n_ints = dh_sur_check_work_len(grid->nz, grid->ny, grid->nx, stride);
    });
}

extern "C" int dh_sur_check_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                const dh_photon_model *model, const float *d_terma,
                                const float *d_density, const float *d_electron, float *d_dose,
                                const float *d_dmax, int32_t stride, float tol, int32_t dilate,
                                int32_t *d_work, int64_t work_len, float *d_plain, void *stream)
{
    return guarded([&] {
        DH_REQUIRE(dilate >= 0 && dilate <= 2, "dilate must be 0, 1 or 2");
        dh_device_enter(gpu_id);
        nvox_of(grid);
        DH_REQUIRE(cp && d_terma && d_density && d_electron && d_dose && d_dmax && d_work && d_plain,
                   "null argument");
        DH_REQUIRE(stride >= 1 && stride <= 64, "lattice stride must be 1..64");
        DH_REQUIRE(work_len >= dh_sur_check_work_len(grid->nz, grid->ny, grid->nx, stride),
                   "work buffer too small (dh_sur_check_work_size)");
        dh_ccc_args ca = make_ccc_args(grid, cp, model);
        dh_pad p = ca.p;
        void *trp = dh_scratch(gpu_id, DH_SLOT_TR_PAD, dh_tr_pad_bytes(ca));
        dh_launch_sur_check(ca, d_terma, d_density, d_electron, d_dose, d_dmax, stride, tol, dilate, d_work,
                            d_plain, trp, (hipStream_t)stream);
    });
}

extern "C" int dh_photon_dose_surrogate_dev(int gpu_id, const dh_sur_net *net, const dh_grid *grid,
                                            const dh_photon_cp *cp, const dh_photon_model *model,
                                            const float *mlc_rows, int32_t n_pairs, const float *d_density,
                                            float *d_dose_out, float *d_terma_out, float *d_electron_out,
                                            int32_t stride, float tol, int32_t dilate, void *stream)
{
    return guarded([&] {
        DH_REQUIRE(net, "null surrogate network");
        DH_REQUIRE(stride >= 1 && stride <= 64, "lattice stride must be 1..64");
        DH_REQUIRE(dilate >= 0 && dilate <= 2, "dilate must be 0, 1 or 2");
        dh_device_enter(gpu_id);
        photon_device(gpu_id, grid, cp, model, mlc_rows, n_pairs, d_density, d_dose_out, d_terma_out,
                      d_electron_out, nullptr, (hipStream_t)stream, net, stride, tol, dilate);
    });
}

extern "C" int dh_sur_dose_dev(const dh_sur_net *net, const dh_grid *grid, const dh_photon_cp *cp,
                               const dh_photon_model *model, const float *d_terma, const float *d_density,
                               const float *d_electron, int32_t stride, float tol, int32_t dilate,
                               float *d_dose_out, void *stream)
{
    return guarded([&] {
        DH_REQUIRE(net && cp && d_terma && d_density && d_electron && d_dose_out, "null argument");
        DH_REQUIRE(stride >= 1 && stride <= 64, "lattice stride must be 1..64");
        DH_REQUIRE(dilate >= 0 && dilate <= 2, "dilate must be 0, 1 or 2");
        dh_device_enter(dh_sur_net_gpu(net));
        nvox_of(grid);
        dh_ccc_args ca = make_ccc_args(grid, cp, model);
        float dir[3];
        sur_direction(cp, dir);
        dh_launch_sur_dose(net, ca, d_terma, d_density, d_electron, dir, stride, tol, dilate, d_dose_out,
                           (hipStream_t)stream);
    });
}

extern "C" int dh_sur_dose_batch_dev(const dh_sur_net *net, const dh_grid *grid, int32_t n_cp,
                                     const dh_photon_cp *cps, const dh_photon_model *model,
                                     const float *const *d_terma, const float *d_density,
                                     const float *const *d_electron, int32_t stride, float tol, int32_t dilate,
                                     float *const *d_dose_out, void *stream)
{
    return guarded([&] {
        DH_REQUIRE(net && cps && d_terma && d_density && d_electron && d_dose_out, "null argument");
        DH_REQUIRE(n_cp >= 1 && n_cp <= 4096, "control point count must be 1..4096");
        DH_REQUIRE(stride >= 1 && stride <= 64, "lattice stride must be 1..64");
        DH_REQUIRE(dilate >= 0 && dilate <= 2, "dilate must be 0, 1 or 2");
        for (int i = 0; i < n_cp; i++)
            DH_REQUIRE(d_terma[i] && d_electron[i] && d_dose_out[i], "null control point volume");
        dh_device_enter(dh_sur_net_gpu(net));
        const size_t bytes = sizeof(float) * nvox_of(grid);
        {
            // outputs are written while other points' inputs are still read
            // (and from two streams): no output may overlap another output
            // or any input volume
            std::vector<std::pair<uintptr_t, int>> r;   // (start, 1 = output)
            for (int i = 0; i < n_cp; i++) {
                r.push_back({(uintptr_t)d_dose_out[i], 1});
                r.push_back({(uintptr_t)d_terma[i], 0});
                r.push_back({(uintptr_t)d_electron[i], 0});
            }
            r.push_back({(uintptr_t)d_density, 0});
            std::sort(r.begin(), r.end());
            uintptr_t end_out = 0, end_any = 0;
            for (const auto &e : r) {
                // an output starting inside any earlier range, or any range
                // starting inside an earlier output, overlaps
                DH_REQUIRE(!(e.second == 1 && e.first < end_any) && !(e.first < end_out),
                           "output volumes must not overlap each other or the input volumes");
                end_any = std::max(end_any, e.first + bytes);
                if (e.second == 1) end_out = std::max(end_out, e.first + bytes);
            }
        }
        std::vector<dh_ccc_args> cas;
        std::vector<float> dirs(3 * (size_t)n_cp);
        cas.reserve(n_cp);
        for (int i = 0; i < n_cp; i++) {
            cas.push_back(make_ccc_args(grid, &cps[i], model));
            sur_direction(&cps[i], &dirs[3 * (size_t)i]);
        }
        dh_launch_sur_dose_batch(net, cas.data(), n_cp, d_terma, d_density, d_electron,
                                 reinterpret_cast<const float (*)[3]>(dirs.data()), stride, tol, dilate, d_dose_out,
                                 (hipStream_t)stream);
    });
}

// ---- enqueue-only entry points for the scheduler ---------------------------
void dh_photon_enqueue(int dev, const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m,
                       const float *mlc_rows, int n_pairs, const float *d_density, float *d_dose,
                       const dh_sur_net *net, int stride, float tol, int dilate, hipStream_t s, bool reuse_wet)
{
    photon_device(dev, grid, cp, m, mlc_rows, n_pairs, d_density, d_dose, nullptr, nullptr, nullptr, s, net,
                  stride, tol, dilate, reuse_wet);
}

void dh_proton_wet_enqueue(int dev, const dh_grid *grid, const dh_proton_beam *beam, const float *d_rlsp,
                           float *d_wet, hipStream_t s)
{
    proton_wet_device(dev, grid, beam, d_rlsp, d_wet, nullptr, s);
}

void dh_proton_plan_enqueue(dh_proton_plan *p, const float *d_wet, float *d_dose, hipStream_t s)
{
    plan_dose(p, d_wet, d_dose, s);
}

int dh_proton_plan_device(const dh_proton_plan *p) { return p->dev; }

extern "C" int dh_autotune_pencil(dh_proton_plan *p, const float *d_wet, int32_t reps, int32_t *best,
                                  float *ms_out, void *stream)
{
    return guarded([&] {
        DH_REQUIRE(p && d_wet, "null argument");
        dh_device_enter(p->dev);
        hipStream_t s = (hipStream_t)stream;
        const size_t n = (size_t)p->g.nx * p->g.ny * p->g.nz;
        float *ref = (float *)dh_scratch(p->dev, DH_SLOT_TMP0, sizeof(float) * n);
        float *out = (float *)dh_scratch(p->dev, DH_SLOT_TMP1, sizeof(float) * n);
        int b = 0;
        float bms = 0.0f;
        for (int c = 0; c < 2 * DH_TILE_CHOICES; c++) {
            float *o = c == 0 ? ref : out;
            const float ms = dh_time_median([&] {
                dh_launch_pencil(p->g, (float)p->beam.vsadx, (float)p->beam.vsady, p->n_layers, p->d_layers,
                                 p->d_luts, p->d_spots, p->fold, d_wet, o, s, c);
            }, reps, s);
            if (ms_out) ms_out[c] = ms;
            DH_REQUIRE(c == 0 || dh_count_diff(ref, out, (int64_t)n, s) == 0,
                       "pencil variants disagree (tile or LUT staging changed the result)");
            if (c == 0 || ms < bms) { b = c; bms = ms; }
        }
        dh_tune_store("pencil", dh_tune_key_geom(p->g), b, bms);
        if (best) *best = b;
    });
}

// Photon control point: ray-trace voxel tile (8 choices) and the cone
// superposition's sampling layout (pairs / quads), both required to give the
// same dose bit for bit.
extern "C" int dh_autotune_photon(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                  const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                                  const float *d_density, int32_t reps, int32_t *best_raytrace,
                                  int32_t *best_layout, float *ms_raytrace, float *ms_ccc, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        hipStream_t s = (hipStream_t)stream;
        const size_t n = nvox_of(grid);
        DH_REQUIRE(cp && d_density, "null argument");
        dh_terma_args ta = make_terma_args(grid, cp, model, mlc_rows, n_pairs);
        dh_ccc_args ca = make_ccc_args(grid, cp, model);
        const dh_pad p = ca.p;
        float4 *vpad = (float4 *)dh_scratch(gpu_id, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
        float *wref = (float *)dh_scratch(gpu_id, DH_SLOT_WET_RAW, sizeof(float) * n);
        float *out = (float *)dh_scratch(gpu_id, DH_SLOT_OUT0, sizeof(float) * n);
        float *dref = (float *)dh_scratch(gpu_id, DH_SLOT_OUT1, sizeof(float) * n);
        float *terma = (float *)dh_scratch(gpu_id, DH_SLOT_TMP0, sizeof(float) * n);
        float *el = (float *)dh_scratch(gpu_id, DH_SLOT_TMP1, sizeof(float) * n);
        void *trp = dh_scratch(gpu_id, DH_SLOT_TR_PAD, dh_tr_pad_bytes(ca));
        int *act = (int *)dh_scratch(gpu_id, DH_SLOT_ACTIVE, sizeof(int) * n);
        int *cnt = (int *)dh_scratch(gpu_id, DH_SLOT_COUNTER, 256);
        dh_launch_pad_quad(d_density, vpad, ca.g.nx, ca.g.ny, ca.g.nz, DH_AIR_DENSITY, s);
        const std::string key = dh_tune_key_geom(ca.g);
        int best = 0;
        float bms = 0.0f;
        for (int c = 0; c < DH_TILE_CHOICES; c++) {
            float *o = c == 0 ? wref : out;
            const float ms = dh_time_median([&] { dh_launch_raytrace(ca.g, vpad, o, s, c); }, reps, s);
            if (ms_raytrace) ms_raytrace[c] = ms;
            DH_REQUIRE(c == 0 || dh_count_diff(wref, out, (int64_t)n, s) == 0,
                       "ray-trace variants disagree (tile choice changed the result)");
            if (c == 0 || ms < bms) { best = c; bms = ms; }
        }
        dh_tune_store("raytrace", key, best, bms);
        if (best_raytrace) *best_raytrace = best;
        dh_launch_terma(ta, wref, terma, el, s);
        best = 0;
        for (int c = 0; c < 2; c++) {
            float *o = c == 0 ? dref : out;
            const float ms = dh_time_median([&] { dh_launch_ccc(ca, terma, d_density, el, trp, act, cnt, o, s, c); },
                                            reps, s);
            if (ms_ccc) ms_ccc[c] = ms;
            DH_REQUIRE(c == 0 || dh_count_diff(dref, out, (int64_t)n, s) == 0,
                       "cone superposition layouts disagree");
            if (c == 0 || ms < bms) { best = c; bms = ms; }
        }
        dh_tune_store("ccc", key, best, bms);
        if (best_layout) *best_layout = best;
    });
}
