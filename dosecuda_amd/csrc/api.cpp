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
    float step = 0.0f;
};
static wet_memo g_wet_memo[64][DH_MAX_WORKSPACES];

// Depth (cm) below which the beam's-eye-view depth is replaced by the 1 mm
// ray trace: the electron-contamination term is non-zero and steep there
// (it reaches zero at the model's fitted dmax); 0.5 cm of margin.
static float bev_exact_depth(const dh_photon_model *m)
{
    return (float)(std::max(0.0, m->electron_fitted_dmax) + 0.5);
}

// Box of the open leaf pairs on the isocentre plane (head frame, mm), widened
// by 15 mm: beyond it the primary transmission, and with it the electron
// term, is negligible (the leaf edges' erf profiles have the primary source
// size, 1 mm, as their width).  An empty box when every pair is closed.
static void bev_field_box(const dh_terma_args &ta, float box[4])
{
    const float ms = (ta.sad - ta.d_mlc) / ta.sad;
    box[0] = box[2] = 1.0f;
    box[1] = box[3] = -1.0f;
    if (ta.n_pairs == 0 || !(ms > 0.0f)) return;
    float x0 = 1e30f, x1 = -1e30f, y0 = 1e30f, y1 = -1e30f;
    for (int i = 0; i < ta.n_pairs; i++) {
        const float4 r = ta.mlc[i];
        x0 = std::min(x0, std::min(r.x, r.y) / ms);
        x1 = std::max(x1, std::max(r.x, r.y) / ms);
        y0 = std::min(y0, std::min(r.z, r.w) / ms);
        y1 = std::max(y1, std::max(r.z, r.w) / ms);
    }
    const float margin = 15.0f;
    box[0] = x0 - margin; box[1] = x1 + margin;
    box[2] = y0 - margin; box[3] = y1 + margin;
}

// TERMA from the beam's-eye-view depth in one pass (dh_launch_bev_terma)
static bool bev_terma(int dev, const dh_terma_args &ta, const float4 *vpad, const dh_photon_model *m, float *terma,
                      float *electron, float *wet_out, hipStream_t s, int table_slot = DH_SLOT_BEV,
                      bool table_ready = false)
{
    float box[4];
    bev_field_box(ta, box);
    return dh_launch_bev_terma(dev, ta, vpad, bev_exact_depth(m), box, terma, electron, wet_out, s, table_slot,
                               table_ready);
}

static void photon_device(int dev, const dh_grid *grid, const dh_photon_cp *cp,
                          const dh_photon_model *m, const float *mlc_rows, int32_t n_pairs,
                          const float *d_density, float *d_dose, float *d_terma,
                          float *d_electron, float *d_wet, hipStream_t s,
                          const dh_sur_net *net = nullptr, int stride = 3, float tol = 0.01f, int dilate = 2,
                          bool reuse_wet = false)
{
    size_t n = nvox_of(grid);
    DH_REQUIRE(cp && d_density && d_dose, "null argument");
    dh_terma_args ta = dh_make_terma_args(grid, cp, m, mlc_rows, n_pairs);
    dh_ccc_args ca = dh_make_ccc_args(grid, cp, m);
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
    // The checked surrogate path takes its radiological depth from fan-line
    // integrals on a beam's-eye-view grid (step 0 here), or where that grid
    // does not apply from a ray trace with one sample per voxel spacing: its
    // TERMA is an input of the network and of the plain check alike, and the
    // result is judged by gamma against the plain CCCS.  The plain path keeps
    // the reference's ray trace with 1 mm samples.
    // The beam's-eye-view depth also depends on the aperture (its exact
    // shallow trace covers the open leaf pairs' box), so a reused one keeps
    // only its fan-line table and runs its lookup again.
    // On the beam's-eye-view path the depth goes straight into TERMA (one
    // pass, written to d_wet only when the caller asked for it).
    float step = net ? 0.0f : 1.0f;
    bool done = false;
    if (!(reuse_wet && !d_wet && memo.vpad == vpad && memo.wet == wet && memo.step == step)) {
        dh_launch_pad_quad(d_density, vpad, ca.g.nx, ca.g.ny, ca.g.nz, DH_AIR_DENSITY, s);  // IMRTClasses.cu:394
        if (step == 0.0f) {
            done = bev_terma(dev, ta, vpad, m, terma, el, d_wet, s);
            if (!done) step = std::max(1.0f, ca.g.s);
        }
        if (step > 0.0f) dh_launch_raytrace(ca.g, vpad, wet, s, -1, nullptr, step);
    } else if (step == 0.0f) {
        DH_REQUIRE(bev_terma(dev, ta, vpad, m, terma, el, nullptr, s, DH_SLOT_BEV, true),
                   "depth table no longer applies");
        done = true;
    }
    memo.vpad = d_wet ? nullptr : vpad;
    memo.wet = d_wet ? nullptr : wet;
    memo.step = step;
    if (!done) dh_launch_terma(ta, wet, terma, el, s);
    if (net) {
        float dir[3];
        sur_direction(cp, dir);
        dh_launch_sur_dose(net, ca, terma, d_density, el, dir, stride, tol, dilate, d_dose, s);
    } else {
        dh_launch_ccc(ca, terma, d_density, el, trp, act, cnt, d_dose, s);
    }
}

extern "C" int dh_photon_terma_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                                   const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                                   const float *d_density, float step_mm, float *d_terma_out,
                                   float *d_electron_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        const size_t n = nvox_of(grid);
        DH_REQUIRE(cp && d_density && d_terma_out && d_electron_out, "null argument");
        DH_REQUIRE(step_mm == 0.0f || (step_mm >= 1.0f && step_mm <= 16.0f),
                   "ray-trace step must be 0 (beam's-eye-view grid) or 1..16 mm");
        hipStream_t s = (hipStream_t)stream;
        dh_terma_args ta = dh_make_terma_args(grid, cp, model, mlc_rows, n_pairs);
        const dh_pad p = dh_make_pad(ta.g.nx, ta.g.ny, ta.g.nz);
        float4 *vpad = (float4 *)dh_scratch(gpu_id, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
        float *wet = (float *)dh_scratch(gpu_id, DH_SLOT_WET_RAW, sizeof(float) * n);
        dh_launch_pad_quad(d_density, vpad, ta.g.nx, ta.g.ny, ta.g.nz, DH_AIR_DENSITY, s);  // IMRTClasses.cu:394
        bool done = false;
        if (step_mm == 0.0f) {
            done = bev_terma(gpu_id, ta, vpad, model, d_terma_out, d_electron_out, nullptr, s);
            if (!done) step_mm = std::max(1.0f, ta.g.s);
        }
        if (!done) {
            dh_launch_raytrace(ta.g, vpad, wet, s, -1, nullptr, step_mm);
            dh_launch_terma(ta, wet, d_terma_out, d_electron_out, s);
        }
        g_wet_memo[gpu_id & 63][dh_get_workspace()] = wet_memo{};   // the padded volume was rewritten
    });
}

extern "C" int dh_photon_terma_batch_dev(int gpu_id, const dh_grid *grid, int32_t n, const dh_photon_cp *cps,
                                         const dh_photon_model *model, const float *const *mlc_rows,
                                         const int32_t *n_pairs, const float *d_density, float step_mm,
                                         float *const *d_terma_out, float *const *d_electron_out, void *stream)
{
    return guarded([&] {
        dh_device_enter(gpu_id);
        DH_REQUIRE(n >= 1 && cps && model && mlc_rows && n_pairs && d_density && d_terma_out && d_electron_out,
                   "null argument");
        DH_REQUIRE(step_mm == 0.0f || (step_mm >= 1.0f && step_mm <= 16.0f),
                   "ray-trace step must be 0 (beam's-eye-view grid) or 1..16 mm");
        const size_t nb = sizeof(float) * nvox_of(grid);
        std::vector<std::pair<uintptr_t, uintptr_t>> outs;
        for (int i = 0; i < n; i++) {
            DH_REQUIRE(d_terma_out[i] && d_electron_out[i], "null output volume");
            outs.push_back({(uintptr_t)d_terma_out[i], (uintptr_t)d_terma_out[i] + nb});
            outs.push_back({(uintptr_t)d_electron_out[i], (uintptr_t)d_electron_out[i] + nb});
        }
        const uintptr_t d0 = (uintptr_t)d_density, d1 = d0 + nb;
        for (size_t a = 0; a < outs.size(); a++) {
            DH_REQUIRE(!(outs[a].first < d1 && d0 < outs[a].second), "output volumes must not overlap the density");
            for (size_t b = a + 1; b < outs.size(); b++)
                DH_REQUIRE(!(outs[a].first < outs[b].second && outs[b].first < outs[a].second),
                           "output volumes must not overlap each other");
        }
        std::vector<dh_terma_args> ta;
        for (int i = 0; i < n; i++) ta.push_back(dh_make_terma_args(grid, &cps[i], model, mlc_rows[i], n_pairs[i]));
        hipStream_t s = (hipStream_t)stream;
        const dh_pad p = dh_make_pad(ta[0].g.nx, ta[0].g.ny, ta[0].g.nz);
        float4 *vpad = (float4 *)dh_scratch(gpu_id, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
        float *wet[2] = {(float *)dh_scratch(gpu_id, DH_SLOT_WET_RAW, nb),
                         n > 1 ? (float *)dh_scratch(gpu_id, DH_SLOT_WET_RAW2, nb) : nullptr};
        // the padded density once; the points alternate between the caller's
        // stream and the side stream, each with its own depth table and volume
        dh_launch_pad_quad(d_density, vpad, ta[0].g.nx, ta[0].g.ny, ta[0].g.nz, DH_AIR_DENSITY, s);  // IMRTClasses.cu:394
        dh_side *side = n > 1 ? &dh_side_of(gpu_id) : nullptr;
        if (side) {
            DH_HIP(hipEventRecord(side->fork, s));
            DH_HIP(hipStreamWaitEvent(side->st, side->fork, 0));
        }
        for (int i = 0; i < n; i++) {
            const int k = side ? (i & 1) : 0;
            hipStream_t st = k ? side->st : s;
            float step = step_mm;
            bool done = false;
            if (step == 0.0f) {
                done = bev_terma(gpu_id, ta[i], vpad, model, d_terma_out[i], d_electron_out[i], nullptr, st,
                                 k ? DH_SLOT_BEV2 : DH_SLOT_BEV);
                if (!done) step = std::max(1.0f, ta[i].g.s);
            }
            if (!done) {
                dh_launch_raytrace(ta[i].g, vpad, wet[k], st, -1, nullptr, step);
                dh_launch_terma(ta[i], wet[k], d_terma_out[i], d_electron_out[i], st);
            }
        }
        if (side) {
            DH_HIP(hipEventRecord(side->join, side->st));
            DH_HIP(hipStreamWaitEvent(s, side->join, 0));
        }
        g_wet_memo[gpu_id & 63][dh_get_workspace()] = wet_memo{};   // the padded volume was rewritten
    });
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
        dh_terma_args ta = dh_make_terma_args(grid, cp, model, mlc_rows, n_pairs);
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
        dh_ccc_args ca = dh_make_ccc_args(grid, cp, model);
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
        dh_ccc_args ca = dh_make_ccc_args(grid, cp, model);
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
        dh_check_dilate_word(dilate);
        dh_device_enter(gpu_id);
        nvox_of(grid);
        DH_REQUIRE(cp && d_terma && d_density && d_electron && d_dose && d_dmax && d_work && d_plain,
                   "null argument");
        DH_REQUIRE(stride >= 1 && stride <= 64, "lattice stride must be 1..64");
        DH_REQUIRE(work_len >= dh_sur_check_work_len(grid->nz, grid->ny, grid->nx, stride),
                   "work buffer too small (dh_sur_check_work_size)");
        dh_ccc_args ca = dh_make_ccc_args(grid, cp, model);
        dh_launch_sur_check(gpu_id, ca, d_terma, d_density, d_electron, d_dose, d_dmax, stride, tol, dilate, d_work,
                            d_plain, (hipStream_t)stream);
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
        dh_check_dilate_word(dilate);
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
        dh_check_dilate_word(dilate);
        dh_device_enter(dh_sur_net_gpu(net));
        nvox_of(grid);
        dh_ccc_args ca = dh_make_ccc_args(grid, cp, model);
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
        dh_check_dilate_word(dilate);
        for (int i = 0; i < n_cp; i++)
            DH_REQUIRE(d_terma[i] && d_electron[i] && d_dose_out[i], "null control point volume");
        dh_device_enter(dh_sur_net_gpu(net));
        const size_t bytes = sizeof(float) * nvox_of(grid);
        {
            // outputs are written while other points' inputs are still
            // read: no output may overlap another output or any input volume
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
            cas.push_back(dh_make_ccc_args(grid, &cps[i], model));
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

void dh_photon_sur_group_enqueue(int dev, const dh_grid *grid, int n, const dh_photon_cp *cps,
                                 const dh_photon_model *m, const float *const *mlc_rows, const int *n_pairs,
                                 const float *d_density, float *const *d_dose, const dh_sur_net *net, int stride,
                                 float tol, int dilate, hipStream_t s, bool reuse_first)
{
    DH_REQUIRE(net && n >= 1 && n <= kSurGroupMax, "surrogate group of 1..8 control points with a network");
    const size_t nv = nvox_of(grid);
    std::vector<dh_terma_args> ta;
    std::vector<dh_ccc_args> cas;
    std::vector<float> dirs(3 * (size_t)n);
    for (int i = 0; i < n; i++) {
        ta.push_back(dh_make_terma_args(grid, &cps[i], m, mlc_rows[i], n_pairs[i]));
        cas.push_back(dh_make_ccc_args(grid, &cps[i], m));
        sur_direction(&cps[i], &dirs[3 * (size_t)i]);
    }
    const dh_pad p = cas[0].p;
    float4 *vpad = (float4 *)dh_scratch(dev, DH_SLOT_VOL_PAD, sizeof(float4) * (size_t)p.pxy * p.pz);
    float *wet = (float *)dh_scratch(dev, DH_SLOT_WET_RAW, sizeof(float) * nv);
    float *terma = (float *)dh_scratch(dev, DH_SLOT_TMP0, sizeof(float) * nv * (size_t)n);
    float *el = (float *)dh_scratch(dev, DH_SLOT_TMP1, sizeof(float) * nv * (size_t)n);
    wet_memo &memo = g_wet_memo[dev & 63][dh_get_workspace()];
    std::vector<const float *> tp(n), ep(n);
    // the padded density once per group (the caller's previous point on this
    // workspace left it there when reuse_first)
    if (!(reuse_first && memo.vpad == vpad))
        dh_launch_pad_quad(d_density, vpad, cas[0].g.nx, cas[0].g.ny, cas[0].g.nz, DH_AIR_DENSITY, s);  // IMRTClasses.cu:394
    for (int i = 0; i < n; i++) {
        // each point's depth as photon_device forms it for the surrogate path
        // (beam's-eye-view grid, else a ray trace at one sample per voxel
        // spacing); when the source is the same as the previous point's on
        // this workspace, its fan-line table is reused
        const bool same = i > 0 ? (cps[i].iso[0] == cps[i - 1].iso[0] && cps[i].iso[1] == cps[i - 1].iso[1] &&
                                   cps[i].iso[2] == cps[i - 1].iso[2] && cps[i].ga == cps[i - 1].ga &&
                                   cps[i].ta == cps[i - 1].ta)
                                : reuse_first;
        float step = 0.0f;
        float *ti = terma + nv * (size_t)i, *ei = el + nv * (size_t)i;
        bool done = false;
        if (!(same && memo.vpad == vpad && memo.wet == wet && memo.step == step)) {
            done = bev_terma(dev, ta[i], vpad, m, ti, ei, nullptr, s);
            if (!done) {
                step = std::max(1.0f, cas[i].g.s);
                dh_launch_raytrace(cas[i].g, vpad, wet, s, -1, nullptr, step);
            }
        } else {
            // same source: the fan-line table stands, the aperture-dependent
            // lookup runs again
            DH_REQUIRE(bev_terma(dev, ta[i], vpad, m, ti, ei, nullptr, s, DH_SLOT_BEV, true),
                       "depth table no longer applies");
            done = true;
        }
        memo.vpad = vpad;
        memo.wet = wet;
        memo.step = step;
        tp[i] = ti;
        ep[i] = ei;
        if (!done) dh_launch_terma(ta[i], wet, ti, ei, s);
    }
    dh_launch_sur_dose_batch(net, cas.data(), n, tp.data(), d_density, ep.data(),
                             reinterpret_cast<const float (*)[3]>(dirs.data()), stride, tol, dilate, d_dose, s);
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
        dh_terma_args ta = dh_make_terma_args(grid, cp, model, mlc_rows, n_pairs);
        dh_ccc_args ca = dh_make_ccc_args(grid, cp, model);
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
