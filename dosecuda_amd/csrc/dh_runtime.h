// dh_runtime.h -- host runtime of the dose engine: errors, per-device scratch
// arena, default streams, launch timing, and the kernel launchers shared by
// the C ABI (api.cpp) and the batching code.
#pragma once

#include "dh_common.h"
#include "../../include/dosehip.h"

#include <functional>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

struct dh_error : std::runtime_error {
    int code;
    dh_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define DH_HIP(expr)                                                          \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) throw dh_error(DH_E_RUNTIME, hipGetErrorString(e_)); \
    } while (0)

#define DH_REQUIRE(cond, msg)                                                 \
    do {                                                                      \
        if (!(cond)) throw dh_error(DH_E_INVALID, (msg));                     \
    } while (0)

// Records the thread-local message returned by dh_last_error().
int dh_set_error(int code, const char *msg);

// Runs f, mapping exceptions to DH_E_* codes and the thread-local message.
template <class F>
static inline int dh_guarded(F &&f)
{
    try {
        f();
        return DH_OK;
    } catch (const dh_error &e) {
        return dh_set_error(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return dh_set_error(DH_E_NOMEM, "Not enough host memory");
    } catch (const std::exception &e) {
        return dh_set_error(DH_E_RUNTIME, e.what());
    }
}

// Selects the device and initialises its state once.
void dh_device_enter(int dev);
// Library stream of a device (non-blocking, default priority).
hipStream_t dh_stream(int dev);
// Grow-only named scratch buffer on a device, in the calling thread's
// current workspace.  Growing synchronises the device before releasing the
// old allocation (so it must not happen during stream capture).  Work that
// uses the same workspace from different streams at the same time races;
// the scheduler gives each of its streams a workspace of its own.
constexpr int DH_SCHED_WORKSPACES = 32;                      // 1..32: the scheduler's streams
constexpr int DH_INFL_STREAMS = 4;                           // streams of dh_influence_count_many
constexpr int DH_WS_INFLUENCE = 1 + DH_SCHED_WORKSPACES;     // their workspaces, 33..36
constexpr int DH_MAX_WORKSPACES = DH_WS_INFLUENCE + DH_INFL_STREAMS;   // 0 = library default
void *dh_scratch(int dev, int slot, size_t bytes);
void dh_set_workspace(int ws);
int dh_get_workspace();

// A second stream per device (default priority) with fork / join events, for
// entry points that spread a batch over the caller's stream and this one.
struct dh_side {
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};
dh_side &dh_side_of(int dev);
int dh_current_device();

enum dh_slot {
    DH_SLOT_VOL_PAD = 0,   // quad-packed padded volume for the ray trace
    DH_SLOT_WET_RAW,       // unsmoothed WET
    DH_SLOT_TR_PAD,        // padded {TERMA, density} pairs for the CCCS
    DH_SLOT_ACTIVE,        // compacted list of CCCS voxels
    DH_SLOT_COUNTER,       // small integer counters
    DH_SLOT_IN0,           // host-buffer staging
    DH_SLOT_IN1,
    DH_SLOT_OUT0,
    DH_SLOT_OUT1,
    DH_SLOT_OUT2,
    DH_SLOT_TMP0,
    DH_SLOT_TMP1,
    DH_SLOT_NN_X,          // surrogate input cells
    DH_SLOT_NN_ACT,        // surrogate activations
    DH_SLOT_NN_Y,          // surrogate fp32 output cells
    DH_SLOT_NN_WORK,       // surrogate check: counters, flags, lists
    DH_SLOT_NN_PLAIN,      // surrogate check: plain values on the lattice
    DH_SLOT_NN_SCAL,       // surrogate scalars (max TERMA, max dose)
    DH_SLOT_CHK_TR,        // surrogate check: per point packed {TERMA, density} and cone partials
    DH_SLOT_BEV,           // beam's-eye-view depth table of the checked surrogate path
    DH_SLOT_BEV2,          // second table and depth volume (batched points on the side stream)
    DH_SLOT_WET_RAW2,
    DH_SLOT_COUNT_
};

// ---------------------------------------------------------------------------
// Tuning table: the variant a launcher uses for a (kernel, shape key), set by
// the autotuner or the caller; launchers fall back to their default choice.
// ---------------------------------------------------------------------------
int dh_tune_lookup(const char *kernel, const std::string &key, int dflt);
void dh_tune_store(const char *kernel, const std::string &key, int choice, double ms);
// "nz,ny,nx,axis" with the dominant axis of the beam direction
std::string dh_tune_key_geom(const dh_geom &g);

// Autotuner helpers: count of differing 32-bit words of two device arrays
// (synchronises s) and the median time of reps runs of f on s (ms).
int64_t dh_count_diff(const void *a, const void *b, int64_t n_words, hipStream_t s);
float dh_time_median(const std::function<void()> &f, int reps, hipStream_t s);
// proton field geometry (no collimator; source at (VSADX + VSADY) / 2)
dh_geom dh_proton_geom(const dh_grid *grid, const dh_proton_beam *beam);

// Voxel tiles of the voxel-parallel kernels: choice 0 = linear runs along x,
// 1.. = bricks (see dh_tile); grid receives the workgroup count.
constexpr int DH_TILE_CHOICES = 12;
dh_tile dh_make_tile(int choice, int nx, int ny, int nz, unsigned *grid);

// Launch timing: when enabled, launches are bracketed by HIP events on the
// launch stream under a kernel name.
struct dh_timed {
    const char *name;
    hipStream_t s;
    void *ev0;
    dh_timed(const char *n, hipStream_t st);
    ~dh_timed();
};

// ---------------------------------------------------------------------------
// Kernel launch interfaces
// ---------------------------------------------------------------------------
constexpr int DH_MAX_PAIRS = 128;    // MLC leaf pairs passed by value
constexpr int DH_MAX_PROFILE = 32;   // off-axis profile points
constexpr int DH_MAX_SPECTRAL = 16;  // spectral energies

// proton layer record (IMPTClasses.cuh:17-26 plus per-layer constants)
struct dh_layer_k {
    int start, n;
    float r80, cut, A, B, C, pad_;
};

struct dh_terma_args {
    dh_geom g;
    int n_pairs, n_prof, n_spec;
    int sorted;                      // open pairs' y edges non-decreasing in the pair index
    float sad, ssd_s, inv_p, inv_s_src, d_mlc, ssw, el_att, esw, el_norm_a, el_norm_b;
    float4 mlc[DH_MAX_PAIRS];        // scaled leaf edges {xL, xR, yB, yT} of open pairs
    float prof_r[DH_MAX_PROFILE], prof_i[DH_MAX_PROFILE], prof_s[DH_MAX_PROFILE];
    float mu_e[DH_MAX_SPECTRAL], pw[DH_MAX_SPECTRAL], sw[DH_MAX_SPECTRAL];
};

// Piecewise-linear HU table (ascending nodes) for the device conversion.
constexpr int DH_HU_TABLE_MAX = 128;
struct dh_hu_table {
    int m;
    double x[DH_HU_TABLE_MAX], y[DH_HU_TABLE_MAX];
};
// out[v] = float32(np.interp(hu[v], t.x, t.y)); hu is float32 or float64
void dh_launch_hu_table(const dh_hu_table &t, const void *hu, bool hu_double, int64_t n, float *out,
                        hipStream_t s);

struct dh_ccc_args {
    dh_geom g;
    dh_pad p;
    float ax, ay, az;               // unit vector iso -> source
    float sp;                       // spacing in cm
    float mu_cal, mu;
    int nsteps[6];
    float L0[6], a[6], b[6], c1[6], c2p[6], bp[6], a2[6];
    float dir[72 * 3];              // cone directions in image space
    float hdir[72 * 3];             // the same in the head frame (before dh_head_to_img)
};

// Cone superposition of up to DH_CCC_MULTI_MAX control points' voxel lists on
// one grid in one launch (k_ccc_multi): the model's row parameters and
// head-frame cone directions once, then per point its geometry, MU and
// buffers.  tr: the point's packed {TERMA, density} (pairs layout, written by
// dh_launch_pack_tr_multi from terma and the shared density); part: 12 floats
// per voxel of the grid; out: the dose of the listed voxels.
constexpr int DH_CCC_MULTI_MAX = 8;
struct dh_ccc_multi {
    dh_pad p;
    float sp, mu_cal;
    int nsteps[6];
    float L0[6], a[6], b[6], c1[6], c2p[6], bp[6], a2[6];
    float hdir[72 * 3];
    uint32_t nvox, cap;
    int n;
    dh_geom g[DH_CCC_MULTI_MAX];
    float mu[DH_CCC_MULTI_MAX];
    const void *tr[DH_CCC_MULTI_MAX];
    const float *terma[DH_CCC_MULTI_MAX];
    const float *electron[DH_CCC_MULTI_MAX];
    const int *list[DH_CCC_MULTI_MAX];
    const int *counter[DH_CCC_MULTI_MAX];
    float *part[DH_CCC_MULTI_MAX];
    float *out[DH_CCC_MULTI_MAX];
};
// the shared fields from cas[0] (every point must share grid and model) and
// each point's geometry and MU; buffers are left null
void dh_ccc_multi_init(dh_ccc_multi &m, const dh_ccc_args *cas, int n);
// bytes of one point's packed volume and of its partial sums
size_t dh_ccc_multi_tr_bytes(const dh_ccc_args &a);
size_t dh_ccc_multi_part_bytes(const dh_ccc_args &a);
void dh_launch_pack_tr_multi(const dh_ccc_multi &m, const float *density, hipStream_t s);
// cone kernel + finish over each point's first *counter[q] list entries (at
// most max_n per point, which sizes the grids)
void dh_launch_ccc_multi(const char *timer, dh_ccc_multi m, int64_t max_n, hipStream_t s);

// Lateral box in the head (beam's-eye-view) frame: voxels whose head x, y lie
// outside [x0, x1] x [y0, y1] are skipped by the ray trace (nothing written)
// and get DH_WET_OUTSIDE from the smoothing.  on = 0: every voxel.
struct dh_hbox {
    float x0, x1, y0, y1;
    int on;
};
constexpr float DH_WET_OUTSIDE = 1e30f;   // cm: beyond every layer's range cut

// quad-packed padded volume of the ray trace (float4 per padded voxel)
void dh_launch_pad_quad(const float *src, float4 *dst, int nx, int ny, int nz, float border,
                        hipStream_t s);
// choice < 0: the tuning table's choice for the shape (default 0); the
// pencil's choice is tile + DH_TILE_CHOICES * (LUT in LDS ? 0 : 1)
// step: sample spacing in mm (1: the reference's ray trace)
void dh_launch_raytrace(const dh_geom &g, const float4 *vol_pad, float *wet, hipStream_t s, int choice = -1,
                        const dh_hbox *box = nullptr, float step = 1.0f);
void dh_launch_smooth(const dh_geom &g, const float *wet_raw, float *wet_out, hipStream_t s, int choice = -1,
                      const dh_hbox *box = nullptr);
void dh_launch_pencil(const dh_geom &g, float vsadx, float vsady, int n_layers,
                      const dh_layer_k *d_layers, const float4 *d_luts, const float *d_spots,
                      bool fold, const float *wet, float *dose, hipStream_t s, int choice = -1);
// The surrogate check's `dilate` word: bits 0-1 the neighbourhood of a
// failed lattice cell that is recomputed too (0 none, 1 the 6 face
// neighbours, 2 all 26), bit 2 the build-up band (every voxel with a
// nonzero electron term is recomputed).
inline void dh_check_dilate_word(int dilate)
{
    DH_REQUIRE(dilate >= 0 && dilate <= 6 && (dilate & 3) != 3, "dilate must be 0, 1 or 2, plus 4 for the build-up band");
}

// TERMA and electron volumes (as dh_launch_terma) from the radiological depth
// (cm) of every voxel by fan-line integrals on a beam's-eye-view grid (one
// density sample per line step of one voxel spacing, trapezoid rule,
// trilinear lookup), the depth formed per voxel and fed straight into TERMA;
// voxels shallower than w_exact_cm whose projection on the isocentre plane
// (source distance a.sad) lies in field_box {x0, x1, y0, y1} (head frame,
// mm; null: none) take the 1 mm ray trace.  The table depends on the
// density, the grid and the source only; table_ready: table_slot already
// holds it for this geometry and density, and only the (aperture-dependent)
// lookup and TERMA run.  wet_out, when not null, receives the depth.  Returns
// false (nothing launched) when the source is not well outside the volume or
// the table would be too large.
bool dh_launch_bev_terma(int dev, const dh_terma_args &a, const float4 *vol_pad, float w_exact_cm,
                         const float *field_box, float *terma, float *electron, float *wet_out, hipStream_t s,
                         int table_slot = DH_SLOT_BEV, bool table_ready = false);
void dh_launch_terma(const dh_terma_args &a, const float *wet, float *terma, float *electron,
                     hipStream_t s);
// Sampling layout of the cone superposition's {TERMA, density} volume:
// 0 = 8-byte pairs, 1 = 32-byte quads with forward x differences (see
// k_pack_tr); choice < 0 takes the
// tuning table's choice ("ccc", default 0).  The scratch for either layout
// is dh_tr_pad_bytes.
int dh_ccc_layout(const dh_ccc_args &a, int choice);
size_t dh_tr_pad_bytes(const dh_ccc_args &a);
void dh_launch_ccc(const dh_ccc_args &a, const float *terma, const float *density,
                   const float *electron, void *tr_pad, int *active, int *counter,
                   float *dose, hipStream_t s, int choice = -1);

// Cone superposition on an explicit list of n voxel indices (each < nx*ny*nz);
// only the listed voxels of `dose` are written.
void dh_launch_ccc_list(const dh_ccc_args &a, const float *terma, const float *density,
                        const float *electron, void *tr_pad, const int *list, int n,
                        int *counter, float *dose, hipStream_t s);

// The padded {TERMA, density} volume in a layout.
void dh_launch_pack_tr(const dh_ccc_args &a, const float *terma, const float *density, void *tr_pad, int layout,
                       hipStream_t s);

// Cone superposition on a voxel list whose length is read on the device from
// *counter (at most max_n); tr_pad already packed in `layout`.
void dh_launch_ccc_counted(const dh_ccc_args &a, const void *tr_pad, int layout, const float *electron,
                           const int *list, const int *counter, int64_t max_n, float *dose, hipStream_t s);

// Checked surrogate: plain CCCS on the lattice, per-cell comparison with the
// surrogate dose, plain CCCS on the voxels of failed and uncovered cells.
// work: int32 [4 + 4 * cells + voxels] (counters, lattice flags, failure
// flags, recompute flags, lattice list, fallback list); plain: fp32 voxels.
int64_t dh_sur_check_work_len(int nz, int ny, int nx, int stride);

// Enqueue-only entry points for the scheduler (no host synchronisation, no
// host copies, no allocation once the calling thread's workspace has grown
// to the job's size): photon dose of one control point (plain CCCS, or the
// checked surrogate when net is given), proton WET, proton dense dose of a
// plan, and the fill pass of a prepared influence matrix.
struct dh_sur_net;
struct dh_proton_plan;
struct dh_influence;
void dh_photon_enqueue(int dev, const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m,
                       const float *mlc_rows, int n_pairs, const float *d_density, float *d_dose,
                       const dh_sur_net *net, int stride, float tol, int dilate, hipStream_t s,
                       bool reuse_wet = false);
// Checked surrogate doses of n (1..kSurGroupMax) control points of one grid,
// density, model and network: each point's depth (reused from the previous
// point when the source is the same; reuse_first: from the caller's previous
// point) and TERMA into the workspace, then one batched network and check.
// Each dose equals dh_photon_enqueue's for that point bit for bit.
constexpr int kSurGroupMax = 8;
void dh_photon_sur_group_enqueue(int dev, const dh_grid *grid, int n, const dh_photon_cp *cps,
                                 const dh_photon_model *m, const float *const *mlc_rows, const int *n_pairs,
                                 const float *d_density, float *const *d_dose, const dh_sur_net *net, int stride,
                                 float tol, int dilate, hipStream_t s, bool reuse_first);
void dh_proton_wet_enqueue(int dev, const dh_grid *grid, const dh_proton_beam *beam, const float *d_rlsp,
                           float *d_wet, hipStream_t s);
void dh_proton_plan_enqueue(dh_proton_plan *p, const float *d_wet, float *d_dose, hipStream_t s);
int dh_proton_plan_device(const dh_proton_plan *p);
// Uploads the fill pass's offsets and blocks once (synchronous); later
// peak/count passes on the same handle invalidate them.
void dh_influence_fill_prepare(dh_influence *f, hipStream_t s);
void dh_influence_fill_enqueue(dh_influence *f, const float *d_wet, uint32_t *d_col, float *d_val, hipStream_t s);
int dh_influence_device(const dh_influence *f);
int64_t dh_influence_nnz(const dh_influence *f);

// Native surrogate network (layer table and device weight pointers).
struct dh_sur_net;
int dh_sur_net_gpu(const dh_sur_net *net);
// Checked surrogate dose of one control point from its TERMA and electron
// volumes (pack, network, finish, lattice check, plain fallback), with the
// workspace of the calling thread; a = the control point's cone arguments.
void dh_launch_sur_dose(const dh_sur_net *net, const dh_ccc_args &a, const float *terma,
                        const float *density, const float *electron, const float dir[3],
                        int stride, float tol, int dilate, float *dose, hipStream_t s);
// The same for ncp control points on one grid and density, the network run
// once per batch of points; each dose equals the single-point one bit for bit.
void dh_launch_sur_dose_batch(const dh_sur_net *net, const dh_ccc_args *cas, int ncp,
                              const float *const *terma, const float *density, const float *const *electron,
                              const float (*dirs)[3], int stride, float tol, int dilate, float *const *dose,
                              hipStream_t s);
// Lattice check and plain fallback of one surrogate dose with the caller's
// work (dh_sur_check_work_len ints) and plain (one float per voxel) buffers.
void dh_launch_sur_check(int dev, const dh_ccc_args &a, const float *terma, const float *density,
                         const float *electron, float *dose, const float *dmax, int stride,
                         float tol, int dilate, int *work, float *plain, hipStream_t s);

// A proton field's spots sorted by energy id (stable), its layers
// (importLayers) and one 400-entry LUT row per layer.  spots holds 8 floats
// per sorted spot: x, y, unit axis direction x/y components, MU; order[q] is
// the caller's row of sorted spot q.
struct dh_proton_tables {
    int n_layers = 0;
    bool fold = true;
    std::vector<dh_layer_k> layers;
    std::vector<float4> luts;
    std::vector<float> spots;
    std::vector<int> order;
};
dh_proton_tables dh_build_proton_tables(const dh_proton_beam *beam, const dh_proton_lut *lut,
                                        const float *spot_rows, int32_t n_spots);

// Host-side beam set-up shared by every entry point (dosemodule.cu:148-172,
// CudaClasses.cu:17-51, IMRTClasses.cu:30-39).
dh_geom dh_make_geom(const dh_grid &grid, const float iso[3], double gantry, double couch,
                     double collimator, float src_dist);

// Photon control-point set-up (host only): geometry, model checks, TERMA and
// cone-superposition kernel arguments.
dh_geom dh_photon_geom(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m);
void dh_check_photon_model(const dh_photon_model *m);
dh_terma_args dh_make_terma_args(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m,
                                 const float *mlc_rows, int32_t n_pairs);
dh_ccc_args dh_make_ccc_args(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m);

// Influence-matrix table bounds of one layer (host only): per-layer maxima
// of the total sigma and the dose factor over WET in [0, 1.1 R80] and source
// distances in [dist_min, dist_max], the depth-resolved bounds per 1 mm cell
// {sigma_t max, P max, sigma_ht max, H max}, and the radius beyond which
// those put every voxel below a threshold.
void dh_infl_layer_bounds(const dh_layer_k &L, const float4 *lut, float dist_min, float dist_max,
                          float &sig_max, float &dose_max);
std::vector<float4> dh_infl_depth_cells(const dh_layer_k &L, const float4 *lut, float dist_min, float dist_max);
float dh_infl_depth_radius(const std::vector<float4> &cells, float thr);
