/*
 * dosehip.h -- C ABI of the MI355X dose engine (libdosehip.so).
 *
 * Plain C types only: pointers, sizes and POD structs.  Every entry point
 * returns DH_OK (0) or a negative DH_E* code; dh_last_error() then holds a
 * thread-local message.  Host-buffer entry points are synchronous (they
 * return when the result is in the caller's buffer).  *_dev entry points take
 * device pointers and a hipStream_t passed as void*; they only enqueue work.
 *
 * Reference interfaces replaced (DoseCUDA/dose_kernels):
 *   dh_proton_raytrace  <- proton_raytrace (dosemodule.cu:104-201)
 *                          + proton_raytrace_cuda (IMPTClasses.cu:274-299)
 *   dh_proton_spot      <- proton_spot (dosemodule.cu:243-374)
 *                          + proton_spot_cuda (IMPTClasses.cu:301-340)
 *   dh_photon_dose      <- photon_dose (dosemodule.cu:377-557)
 *                          + photon_dose_cuda (IMRTClasses.cu:353-405)
 * The Python module dose_kernels (dosemodule.cu:560-595) becomes the ctypes
 * module dosecuda_amd/dose_kernels.py, which binds exactly these symbols.
 */
#ifndef DOSEHIP_H
#define DOSEHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DH_OK 0
#define DH_E_INVALID (-1)   /* bad argument (shape, size, null pointer)   */
#define DH_E_RUNTIME (-2)   /* HIP runtime error (message from HIP)        */
#define DH_E_NOMEM   (-3)   /* host allocation failed                      */

#define DH_LUT_LENGTH 400   /* IMPTClasses.cuh:6                           */
#define DH_AIR_DENSITY 0.0012f /* IMRTClasses.cuh:6                        */

/* Grid as the caller holds it: NumPy C-order (nz, ny, nx) float32 volume,
 * isotropic spacing spacing[0] mm (dosemodule.cu:146), origin in mm. */
typedef struct dh_grid {
    int32_t nz, ny, nx;
    float spacing;
    float origin[3];
} dh_grid;

/* Proton field (dosemodule.cu:114-141): raw caller angles in degrees;
 * the +180 deg gantry shift and iso - origin are applied inside. */
typedef struct dh_proton_beam {
    float iso[3];
    double gantry_angle;
    double couch_angle;
    double vsadx, vsady;     /* model.VSADX / VSADY, mm */
} dh_proton_beam;

/* Proton beam-model tables (dosemodule.cu:260-266,340-349): row-major
 * [n_energies][row_stride]; the first DH_LUT_LENGTH columns are used.
 * divergence_params is [n_energies][5] = energy, R80, A, B, C. */
typedef struct dh_proton_lut {
    int32_t n_energies;
    int32_t row_stride;
    const float *divergence_params;
    const float *depths;
    const float *sigmas;
    const float *idds;
} dh_proton_lut;

/* Photon control point (dosemodule.cu:449-466): raw angles in degrees. */
typedef struct dh_photon_cp {
    float iso[3];
    double mu, ga, ca, ta;
} dh_photon_cp;

/* Photon beam model (dosemodule.cu:386-437, IMRTClasses.cuh:23-57).
 * kernel is [6 zenith rows][theta_deg, Am, am, Bm, bm, ray_length_cm]. */
typedef struct dh_photon_model {
    int32_t n_profile;
    const float *profile_radius, *profile_intensities, *profile_softening;
    int32_t n_spectral;
    const float *spectrum_attenuation_coefficients;
    const float *spectrum_primary_weights;
    const float *spectrum_scatter_weights;
    double mu_calibration, primary_source_distance, scatter_source_distance;
    double primary_source_size, scatter_source_size, mlc_distance;
    double scatter_source_weight, electron_attenuation, electron_source_weight;
    double electron_fitted_dmax, jaw_transmission, mlc_transmission;
    int32_t has_xjaws, has_yjaws;
    float kernel[36];
} dh_photon_model;

/* ---- library ---------------------------------------------------------- */
const char *dh_last_error(void);
int dh_version(void);
int dh_device_count(int *count);
/* Selects and initialises a device (stream, scratch arena).  Called
 * implicitly by every entry point with its gpu_id. */
int dh_init(int gpu_id);
/* Releases every device buffer the library holds on gpu_id. */
int dh_release(int gpu_id);

/* ---- drop-in, host buffers -------------------------------------------- */
/* Smoothed WET (cm), shape of rlsp (IMPTClasses.cu:274-299). */
int dh_proton_raytrace(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                       const float *rlsp, float *wet_out);
/* Pencil-beam dose on the WET grid; spot_rows is [n_spots][4] =
 * x mm, y mm, MU, energy id (dosemodule.cu:211-225). */
int dh_proton_spot(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                   const dh_proton_lut *lut, const float *spot_rows, int32_t n_spots,
                   const float *wet, float *dose_out);
/* CCCS dose before the output factor; mlc_rows is [n_pairs][4] =
 * x1, x2, y_offset, y_width mm at iso.  terma_out / electron_out are
 * optional (NULL) extra outputs the reference keeps on the device. */
int dh_photon_dose(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                   const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                   const float *density, float *dose_out, float *terma_out,
                   float *electron_out);

/* ---- device buffers, stream-ordered ---------------------------------- */
/* Same computations on device pointers.  wet_raw_out (optional) receives
 * the unsmoothed WET. */
int dh_proton_raytrace_dev(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                           const float *d_rlsp, float *d_wet_out, float *d_wet_raw_out,
                           void *stream);
/* dh_proton_raytrace_dev (smoothed WET into d_wet_out) that allocates no
 * device memory on any call.  No reference counterpart: it is for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The padded volume and the unsmoothed
 * WET go to one fixed 64 MiB device array per GPU (held from the library's
 * first use on that GPU until the process ends), so the volume must satisfy
 * (nx+3)(ny+3)(nz+3)*16 + nx*ny*nz*4 <= 64 MiB (each term rounded up to 256
 * bytes; about 145^3 voxels), else DH_E_INVALID.  All calls on one GPU share
 * the array: calls that overlap on one GPU race (order them on one stream).
 * Same kernels, so the WET equals dh_proton_raytrace_dev's bit for bit. */
int dh_proton_raytrace_dev_noalloc(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                                   const float *d_rlsp, float *d_wet_out, void *stream);
/* A proton plan holds one field's sorted spots, layers and per-layer LUT rows
 * on the device (make_spot_array + importLayers, dosemodule.cu:211-225 and
 * IMPTClasses.cu:29-67), so repeated dose evaluations upload nothing. */
typedef struct dh_proton_plan dh_proton_plan;
int dh_proton_plan_create(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                          const dh_proton_lut *lut, const float *spot_rows, int32_t n_spots,
                          dh_proton_plan **out);
/* Layer count, spots per layer and range cut 1.1*R80 (mm) per layer;
 * arrays may be NULL or must hold n_energies entries. */
int dh_proton_plan_layers(const dh_proton_plan *plan, int32_t *n_layers,
                          int32_t *layer_spots, float *layer_cut);
int dh_proton_plan_dose_dev(dh_proton_plan *plan, const float *d_wet, float *d_dose_out,
                            void *stream);
int dh_proton_plan_destroy(dh_proton_plan *plan);
int dh_photon_dose_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                       const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                       const float *d_density, float *d_dose_out, float *d_terma_out,
                       float *d_electron_out, float *d_wet_out, void *stream);

/* TERMA and electron-contamination volumes of a control point from the
 * density (ray trace + termaKernel, IMRTClasses.cu:209-250,353-405).  The
 * radiological depth: step_mm 1..16, a ray trace sampled every step_mm (1:
 * the reference's 1 mm samples); step_mm 0, the checked surrogate path's
 * depth, fan-line integrals from the source on a beam's-eye-view grid (one
 * density sample per line per voxel spacing, trapezoid rule) interpolated
 * at each voxel, judged by gamma against the plain CCCS (where the source is
 * not well outside the volume it falls back to the ray trace at one sample
 * per voxel spacing). */
int dh_photon_terma_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                        const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                        const float *d_density, float step_mm, float *d_terma_out,
                        float *d_electron_out, void *stream);

/* dh_photon_terma_dev for n control points of one grid and density (the
 * per-point TERMA of IMRTClasses.cu:353-405, batched; no reference
 * counterpart as a batch): the density is padded once and the points
 * alternate between the caller's stream and a side stream, each with its own
 * depth buffers, so their depth and TERMA launches overlap.  mlc_rows[i] /
 * n_pairs[i] are point i's aperture.  Each output equals dh_photon_terma_dev
 * of that point bit for bit.  Output volumes may not overlap each other or
 * the density. */
int dh_photon_terma_batch_dev(int gpu_id, const dh_grid *grid, int32_t n, const dh_photon_cp *cps,
                              const dh_photon_model *model, const float *const *mlc_rows,
                              const int32_t *n_pairs, const float *d_density, float step_mm,
                              float *const *d_terma_out, float *const *d_electron_out, void *stream);

/* ---- batched proton dose-influence matrix (CSR) ---------------------- */
/* No reference counterpart: the reference only sums weighted spots
 * (IMPTClasses.cu:204-271).  Row q of the matrix is sorted spot q (energy-id
 * order, stable; dh_influence_info gives the caller's spot row) with MU = 1;
 * columns are voxel indices i + nx*(j + ny*k) in ascending order; values are
 * the pencil-beam dose of that spot, kept where >= rel_threshold * row peak.
 * spots_per_block (4, 8 or 16) and row_chunks (1..256) are tuning knobs.
 * Usage: create -> count (synchronises; gives nnz and row_ptr, whose entries
 * start at `base` so fields can share one CSR) -> caller allocates col/val ->
 * fill (stream-ordered, with the same WET as count) -> destroy. */
typedef struct dh_influence dh_influence;
int dh_influence_create(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                        const dh_proton_lut *lut, const float *spot_rows, int32_t n_spots,
                        int32_t spots_per_block, int32_t row_chunks, float rel_threshold,
                        dh_influence **out);
/* The field's smoothed WET (as dh_proton_raytrace_dev) computed only where
 * the matrix can have entries: inside a beam's-eye-view box around the spot
 * axes, widened per layer by the radius beyond which the layer's table
 * bounds put every voxel below an estimated threshold (times margin_scale,
 * 1 by default); voxels outside get 1e30 cm (beyond every range cut).
 * dh_influence_count with this buffer checks every spot's radius at its
 * actual threshold against the box and, when one does not fit, recomputes
 * the WET on every voxel into the same buffer first, so the matrix always
 * equals the one built from the full-grid WET.  No reference counterpart
 * (the reference traces every voxel, IMPTClasses.cu:274-299). */
int dh_influence_wet(dh_influence *f, const float *d_rlsp, float margin_scale, float *d_wet_out,
                     void *stream);
/* Footprint box {x0, x1, y0, y1} (head frame, mm) of the last
 * dh_influence_wet, whether it is still in use (0 after a fallback) and the
 * number of fallbacks to the full-grid WET so far. */
int dh_influence_footprint(const dh_influence *f, float box[4], int32_t *active, int64_t *fallbacks);
int dh_influence_count(dh_influence *f, const float *d_wet, int64_t base, int64_t *nnz,
                       int64_t *row_ptr, void *stream);
/* dh_influence_wet (footprint WET from the shared d_rlsp into d_wet[q]) and
 * dh_influence_count for n fields of one device at once: the fields' passes
 * run on library streams of their own (four, with scratch workspaces of
 * their own), overlapping on the GPU, after the work already on `stream`,
 * with one host synchronisation for all.  Field q's entries start at base +
 * nnz[0] + ... + nnz[q - 1]; nnz[q] and row_ptr[q] (n_rows + 1 entries,
 * skipped when row_ptr or row_ptr[q] is null) equal what the single-field
 * calls in that order give.  Fields and WET buffers must be distinct.  No
 * reference counterpart (the reference builds no influence matrix). */
int dh_influence_count_many(dh_influence *const *fields, int32_t n, const float *d_rlsp, float margin_scale,
                            float *const *d_wet, int64_t base, int64_t *nnz, int64_t *const *row_ptr,
                            void *stream);
int dh_influence_fill(dh_influence *f, const float *d_wet, uint32_t *d_col, float *d_val,
                      void *stream);
int dh_influence_info(const dh_influence *f, int32_t *n_rows, int32_t *row_spot, float *row_peak);
/* Sweep counters of the last peak/count/fill passes (4 each: lane visits,
 * in-range visits, 64-voxel chunks, rows) and each row's truncation radius. */
int dh_influence_stats(const dh_influence *f, uint64_t *stats12, float *row_radius);
int dh_influence_destroy(dh_influence *f);

/* ---- HU conversion --------------------------------------------------- */
/* Replaces the host conversions RLSPFromHU (plan_impt.py:112-122) and
 * DensityFromHU (plan_imrt.py:114-124): d_out[v] = float32(np.interp(
 * d_hu[v], xp, fp)) over the machine's HU table (m <= 128 ascending nodes;
 * d_hu float32, or float64 when hu_is_double), evaluated on the device with
 * np.interp's FP64 arithmetic, so the result equals the host conversion bit
 * for bit. */
int dh_hu_convert_dev(int gpu_id, const void *d_hu, int32_t hu_is_double, int64_t n, const double *xp,
                      const double *fp, int32_t m, float *d_out, void *stream);

/* ---- single stages on device pointers (stage-wise parity) ------------ */
/* Unsmoothed WET of a volume sampled with the given border value, source
 * at src_dist from iso (CudaClasses.cu:180-228). */
int dh_stage_raytrace_dev(int gpu_id, const dh_grid *grid, const float iso[3],
                          double gantry, double couch, float src_dist, float border,
                          const float *d_vol, float *d_wet_out, void *stream);
/* Lateral WET smoothing of a given WET (IMPTClasses.cu:139-202). */
int dh_stage_smooth_dev(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                        const float *d_wet_raw, float *d_wet_out, void *stream);
/* TERMA and electron volumes from a given WET (IMRTClasses.cu:209-250). */
int dh_stage_terma_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                       const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                       const float *d_wet, float *d_terma_out, float *d_electron_out,
                       void *stream);
/* Cone superposition from given TERMA, density and electron volumes
 * (IMRTClasses.cu:252-350). */
int dh_stage_ccc_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                     const dh_photon_model *model, const float *d_terma,
                     const float *d_density, const float *d_electron, float *d_dose_out,
                     void *stream);

/* Cone superposition for an explicit list of voxel indices (each in
 * [0, nx*ny*nz)); only listed voxels of d_dose_out are written.  Used as the
 * per-voxel plain fallback of the surrogate. */
int dh_stage_ccc_list_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                          const dh_photon_model *model, const float *d_terma,
                          const float *d_density, const float *d_electron,
                          const int32_t *d_list, int32_t n, float *d_dose_out, void *stream);

/* ---- learned CCCS surrogate: BF16 MFMA building blocks ---------------- */
/* No reference counterpart.  Tensors are channels-last bf16 [n][d][h][w][c].
 * Convolution (k = 1 or 3, stride 1, zero "same" padding) of the channel
 * concatenation [A, B]: A has ca channels and, when up_a = 1, half the
 * spatial size (read nearest-upsampled); B (skip input, cb channels, may be
 * NULL with cb = 0) has the output's size.  Weights in the
 * [taps][cout_pad][ca + cb] bf16 layout, FP32 accumulation, optional bias
 * and LeakyReLU (act = 1); writes bf16 y and/or fp32 y32.  ca, cb multiples
 * of 16, cout_pad of 32.  nco (output tiles of 32 channels per wave: 1, 2, 4)
 * and ksplit (waves splitting K per workgroup: 1..16) pick the kernel
 * variant; ksplit -1 / -2 select the LDS brick kernel (4- / 2-row bricks,
 * nco 1 or 2), ksplit -3 the 16-output-channel kernels (nco = 4 or 8 brick
 * rows for k = 3, nco = 1 for k = 1); 0, 0 lets the library choose. */
int dh_nn_conv3d_dev(const void *xa, int32_t ca, int32_t up_a, const void *xb, int32_t cb,
                     int32_t n, int32_t d, int32_t h, int32_t w, const void *wt,
                     const float *bias, int32_t cout, int32_t cout_pad, int32_t ksize,
                     int32_t act, float slope, void *y, float *y32, int32_t nco,
                     int32_t ksplit, void *stream);
/* 2x2x2 average pooling of channels-last bf16 (even d, h, w). */
int dh_nn_pool2_dev(const void *x, int32_t n, int32_t d, int32_t h, int32_t w, int32_t c,
                    void *y, void *stream);
/* dw[taps][cin][cout] += sum_v x[v + off][ci] * dy[v][co] (FP32 atomics). */
int dh_nn_conv3d_wgrad_dev(const void *x, const void *dy, int32_t n, int32_t d, int32_t h,
                           int32_t w, int32_t cin, int32_t cout, int32_t ksize, float *dw,
                           void *stream);
/* fp32 [taps][cin][cout] -> bf16 [taps][rows_pad][cin] (mode 0, forward) or
 * the tap-reversed [taps][rows_pad][cout] (mode 1, input gradient). */
int dh_nn_prep_weights_dev(const float *w, int32_t taps, int32_t cin, int32_t cout, int32_t mode,
                           int32_t rows_pad, void *out, void *stream);
/* Surrogate input [nz'/2][ny'/2][nx'/2][32] bf16 (n' = n rounded up to a
 * multiple of 16): channels 0-7 TERMA / max TERMA and 8-15 density of the
 * 2x2x2 sub-voxels (sub = 4 dz + 2 dy + dx; TERMA 0 and air density beyond
 * the volume), 16-18 the beam direction dir, 19 spacing / 2.5 mm, 20-31
 * zero.  *tmax_out (device) receives max TERMA. */
int dh_sur_pack_dev(int32_t nz, int32_t ny, int32_t nx, const float *d_terma,
                    const float *d_density, const float dir[3], float spacing, void *d_x_out,
                    float *d_tmax_out, void *stream);
/* Surrogate dose from the network output y32 [cells][16] (channels 0-7 the
 * sub-voxels): dose = (y * tmax * out_scale + electron) * mu_scale where
 * TERMA > 0.01 and that value is positive, else 0; *dmax_out (device)
 * receives the dose maximum. */
int dh_sur_finish_dev(int32_t nz, int32_t ny, int32_t nx, const float *d_y32, const float *d_tmax,
                      float out_scale, const float *d_terma, const float *d_electron,
                      float mu_scale, float *d_dose_out, float *d_dmax_out, void *stream);
/* Plain-CCCS check and fallback of a surrogate dose, all on the device: each
 * lattice cell of stride^3 voxels is checked at one voxel, its low corner
 * (every stride-th voxel per axis) when that has TERMA > 0.01, else its
 * centre voxel (offset stride / 2 per axis) when that lies in the volume with
 * TERMA > 0.01; the plain CCCS runs on those voxels and a cell whose
 * surrogate value misses the plain one by more than tol * dmax fails; voxels
 * with TERMA > 0.01 in a failed cell or next to one (dilate & 3: 0 none, 1
 * the 6 face neighbours, 2 all 26 neighbours of cells), or in a cell with no
 * check voxel, get the plain CCCS value in d_dose; with dilate & 4 so do all voxels with TERMA > 0.01
 * whose electron-contamination term is nonzero (the build-up band, WET below
 * the model's electron_fitted_dmax).  Every dilate argument of the surrogate
 * entry points below takes these values (0-2, 4-6).  d_work holds int32
 * [work_len] (dh_sur_check_work_size); afterwards d_work[0] is the number of
 * lattice voxels and d_work[1] the number of fallback voxels.  d_plain:
 * fp32 voxels of scratch. */
int dh_sur_check_work_size(const dh_grid *grid, int32_t stride, int64_t *n_ints);
int dh_sur_check_dev(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                     const dh_photon_model *model, const float *d_terma, const float *d_density,
                     const float *d_electron, float *d_dose, const float *d_dmax, int32_t stride,
                     float tol, int32_t dilate, int32_t *d_work, int64_t work_len, float *d_plain,
                     void *stream);
/* Native surrogate network: the U-Net's layer weights on the device in the
 * forward bf16 layout ([taps][cout_pad][cin], dh_nn_prep_weights_dev mode 0)
 * and fp32 biases, in the order e0.0 e0.1 e1.0 e1.1 e2.0 e2.1 bt.0 bt.1
 * d2.0 d2.1 d1.0 d1.1 d0.0 d0.1 head (widths w0..w3, in_ch input channels,
 * head 1x1 with 16 outputs).  The buffers stay owned by the caller and must
 * outlive the handle. */
typedef struct dh_sur_net dh_sur_net;
int dh_sur_net_create(int gpu_id, const int32_t widths[4], int32_t in_ch,
                      const void *const *d_weights, const float *const *d_biases, float slope,
                      float out_scale, dh_sur_net **out);
int dh_sur_net_destroy(dh_sur_net *net);
/* Forward pass on cz x cy x cx cells (multiples of 8): x [cells][in_ch] bf16
 * -> y32 [cells][16] fp32; activations in the library's scratch. */
int dh_sur_net_forward_dev(const dh_sur_net *net, const void *d_x, int32_t cz, int32_t cy,
                           int32_t cx, float *d_y32, void *stream);
/* Checked surrogate dose of one control point from the plain path's TERMA
 * and electron volumes: pack, network, finish, lattice check and plain
 * fallback (dh_sur_check_dev semantics), all on the device. */
int dh_sur_dose_dev(const dh_sur_net *net, const dh_grid *grid, const dh_photon_cp *cp,
                    const dh_photon_model *model, const float *d_terma, const float *d_density,
                    const float *d_electron, int32_t stride, float tol, int32_t dilate,
                    float *d_dose_out, void *stream);
/* dh_sur_dose_dev for n_cp control points of one grid and density (per-point
 * TERMA, electron and output volumes, arrays of device pointers): the points
 * share network launches, up to 8 per batch, and each dose equals the
 * single-point call's bit for bit.  Output volumes must not overlap one
 * another or any input volume (DH_E_INVALID).  No reference counterpart (the reference
 * runs one control point per photon_dose_cuda call, plan_imrt.py:150-153). */
int dh_sur_dose_batch_dev(const dh_sur_net *net, const dh_grid *grid, int32_t n_cp,
                          const dh_photon_cp *cps, const dh_photon_model *model,
                          const float *const *d_terma, const float *d_density,
                          const float *const *d_electron, int32_t stride, float tol, int32_t dilate,
                          float *const *d_dose_out, void *stream);
/* Lattice (check voxel) and fallback voxel counts of the last checked
 * surrogate dose run by dh_sur_dose_dev / dh_photon_dose_surrogate_dev from
 * this thread on gpu_id (synchronises the device).  After
 * dh_sur_dose_batch_dev they are those of the batch's last point. */
int dh_sur_counts(int gpu_id, int32_t *lattice, int32_t *fallback);
/* dh_photon_dose_dev with the cone superposition replaced by the checked
 * surrogate; the radiological depth comes from the beam's-eye-view grid
 * (dh_photon_terma_dev at step_mm 0), TERMA as in the plain path. */
int dh_photon_dose_surrogate_dev(int gpu_id, const dh_sur_net *net, const dh_grid *grid,
                                 const dh_photon_cp *cp, const dh_photon_model *model,
                                 const float *mlc_rows, int32_t n_pairs, const float *d_density,
                                 float *d_dose_out, float *d_terma_out, float *d_electron_out,
                                 int32_t stride, float tol, int32_t dilate, void *stream);
/* Global gamma index on the voxel grid (Low et al. 1998): out = gamma for
 * voxels with ref >= cut, -1 elsewhere; search radius in voxels. */
int dh_gamma_dev(const float *eval, const float *ref, int32_t nz, int32_t ny, int32_t nx,
                 float spacing, float dta_mm, float dd_abs, float cut, int32_t radius_vox,
                 float *out, void *stream);

/* ---- batching scheduler ----------------------------------------------- */
/* No reference counterpart (DoseCUDA runs one synchronous call per beam,
 * dosemodule.cu:95-538).  A case is one phantom: grid, input volume on the
 * device (density for photon jobs, RLSP for proton jobs) and an optional
 * dose output that every run zeroes and then accumulates the case's photon
 * and proton-dose jobs into, in the order they were added.  The jobs of a
 * case run in order on one of the scheduler's n_streams streams (each with
 * its own scratch workspace); cases go longest first to the least-loaded
 * stream, by per-job GPU times measured in the last eager run (estimated
 * from voxel counts before).  Host arrays (MLC, beam model) are copied when
 * a job is added; device buffers, plans, influence handles and surrogate
 * networks are borrowed and must outlive the scheduler's runs. */
typedef struct dh_sched dh_sched;
int dh_sched_create(int gpu_id, int32_t n_streams, dh_sched **out);
int dh_sched_destroy(dh_sched *sched);
int dh_sched_add_case(dh_sched *sched, const dh_grid *grid, const float *d_volume, float *d_dose,
                      int32_t *case_id);
/* One photon control point: plain CCCS (net = NULL) or the checked
 * surrogate (dh_photon_dose_surrogate_dev). */
int dh_sched_add_photon(dh_sched *sched, int32_t case_id, const dh_photon_cp *cp,
                        const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                        const dh_sur_net *net, int32_t stride, float tol, int32_t dilate);
/* As dh_sched_add_photon, with the control point's dose multiplied by scale
 * (rounded to float32) before it is added to the case dose: a case's jobs
 * then sum Σ dose_cp * scale_cp in job order, as the plan loop does with the
 * output factors (plan_imrt.py:140-153). No reference counterpart (it batches
 * the plan loop). */
int dh_sched_add_photon_scaled(dh_sched *sched, int32_t case_id, const dh_photon_cp *cp,
                               const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                               const dh_sur_net *net, int32_t stride, float tol, int32_t dilate,
                               float scale);
/* Smoothed WET of the case's RLSP for one proton field into d_wet_out. */
int dh_sched_add_proton_wet(dh_sched *sched, int32_t case_id, const dh_proton_beam *beam,
                            float *d_wet_out);
/* Dense dose of a proton plan from a WET, accumulated into the case dose. */
int dh_sched_add_proton_dose(dh_sched *sched, int32_t case_id, dh_proton_plan *plan,
                             const float *d_wet);
/* Fill pass of a counted influence matrix (uploads its fill blocks now;
 * no peak/count pass may run on the handle afterwards). */
int dh_sched_add_influence_fill(dh_sched *sched, int32_t case_id, dh_influence *f,
                                const float *d_wet, uint32_t *d_col, float *d_val);
/* mode 0: eager run with per-job timing (synchronises); mode 1: one HIP
 * graph of the whole batch, captured on first use after an eager warm-up and
 * replayed afterwards (asynchronous).  The batch starts after the work
 * queued on `stream`, whose later work waits for it. */
int dh_sched_run(dh_sched *sched, int32_t mode, void *stream);
/* Case and job counts, the last eager batch time, the stream of each case
 * and each job's last measured time (arrays may be NULL). */
int dh_sched_info(const dh_sched *sched, int32_t *n_cases, int32_t *n_jobs, double *last_ms,
                  int32_t *case_stream, double *job_ms);

/* ---- autotuner --------------------------------------------------------- */
/* No reference counterpart (the reference's launch shapes are compile-time
 * constants, CudaClasses.cuh:14).  The tuning table maps (kernel, shape key)
 * to the variant its launcher uses; the tuners time every variant on the
 * caller's data (median of reps), require the outputs to agree (bit for bit
 * for the voxel-tile and LUT-staging variants, to 1e-5 of the largest value
 * for the convolution's summation orders) and store the fastest.  Keys:
 * "raytrace" / "smooth" / "pencil" / "ccc": "nz,ny,nx,axis" (dominant axis
 * of the beam direction); "conv3d": "n,d,h,w,ca,cb,up,cout,ksize". */
int dh_tune_set(const char *kernel, const char *key, int32_t choice, double ms);
int dh_tune_get(const char *kernel, const char *key, int32_t *choice);   /* -1: none */
int dh_tune_count(int32_t *n);
int dh_tune_entry(int32_t index, char *kernel, int32_t kernel_cap, char *key, int32_t key_cap,
                  int32_t *choice, double *ms);
int dh_tune_clear(void);
/* Voxel tile codes of the ray trace, smoothing and pencil launches (12):
 * 0 linear runs along x; 1-7 bricks 64x4x1, 32x8x1, 16x16x1, 32x4x2, 16x4x4,
 * 8x8x4, 64x2x2 numbered x fastest; 8-11 bricks 16x16x1, 16x4x4, 8x8x4,
 * 32x4x2 numbered y fastest.  Every code gives the same result bit for bit.
 * Ray-trace and smoothing voxel tiles (12 choices each) for a proton field;
 * ms arrays hold 12 entries. */
int dh_autotune_proton_wet(int gpu_id, const dh_grid *grid, const dh_proton_beam *beam,
                           const float *d_rlsp, int32_t reps, int32_t *best_raytrace,
                           int32_t *best_smooth, float *ms_raytrace, float *ms_smooth, void *stream);
/* Photon control point: ray-trace voxel tile (12 choices) and the cone
 * superposition's {TERMA, density} sampling layout (0 float2 pairs, 1 float4
 * quads); ms_raytrace holds 12 entries, ms_ccc 2. */
int dh_autotune_photon(int gpu_id, const dh_grid *grid, const dh_photon_cp *cp,
                       const dh_photon_model *model, const float *mlc_rows, int32_t n_pairs,
                       const float *d_density, int32_t reps, int32_t *best_raytrace,
                       int32_t *best_layout, float *ms_raytrace, float *ms_ccc, void *stream);
/* Pencil-beam voxel tile x LUT staging (LDS or global): 24 choices (tile +
 * 12 when the LUT is read from global memory). */
int dh_autotune_pencil(dh_proton_plan *plan, const float *d_wet, int32_t reps, int32_t *best,
                       float *ms, void *stream);
/* Convolution variant of one layer shape (codes: 1000 + nco for the LDS
 * brick kernel, 1100 + nco for its two-row bricks, 3004 / 3008 for the
 * 16-output-channel brick kernel (4 / 8 rows), 3001 for the 16-channel 1x1
 * kernel, nco * 100 + ksplit otherwise; a rejected variant's code is reported
 * negated); up to 24 variants.  dh_nn_conv3d_dev selects the same kernels
 * with (nco, ksplit) = (nco, -1), (nco, -2), (rows, -3) or (1, -3). */
int dh_autotune_conv3d(int gpu_id, int32_t n, int32_t d, int32_t h, int32_t w, int32_t ca,
                       int32_t cb, int32_t up, int32_t cout, int32_t ksize, int32_t reps,
                       int32_t *best_code, int32_t *codes, float *ms, int32_t *n_tried, void *stream);

/* ---- kernel timing (HIP events on the launch stream) ------------------ */
/* When enabled, every kernel launch of the library is bracketed by HIP
 * events; dh_prof_get sums the elapsed times of the launches of one kernel
 * name since the last reset (synchronising the device first). */
int dh_prof_enable(int on);
int dh_prof_reset(void);
int dh_prof_get(const char *kernel, double *total_ms, int64_t *launches);

#ifdef __cplusplus
}
#endif
#endif /* DOSEHIP_H */
