// host_tables.cpp -- the GPU-free host set-up of the engine: beam geometry,
// proton spot/layer tables, TERMA and cone-superposition kernel arguments
// (with the MLC rectangle merge), and the influence matrix's table bounds.
// No device code and no HIP runtime calls, so the same sources also build
// into a CPU-only sanitizer check (sanitize_host.cpp).
#include "dh_runtime.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

// Host-side beam set-up shared by every entry point (dosemodule.cu:148-172,
// CudaClasses.cu:17-51, IMRTClasses.cu:30-39).
dh_geom dh_make_geom(const dh_grid &grid, const float iso[3], double gantry, double couch,
                     double collimator, float src_dist)
{
    dh_geom g;
    g.nx = grid.nx; g.ny = grid.ny; g.nz = grid.nz;
    g.s = grid.spacing;                 // CudaDose(dims, spacing), dosemodule.cu:146
    g.inv_s = 1.0f / grid.spacing;      // CudaClasses.cuh:101
    for (int a = 0; a < 3; a++) g.iso[a] = iso[a] - grid.origin[a];   // dosemodule.cu:160-164
    float adj = fmodf((float)(gantry + 180.0), 360.0f);               // dosemodule.cu:152
    float ga = adj * DH_PI_F / 180.0f;                                  // CudaClasses.cu:27
    float ta = (float)couch * DH_PI_F / 180.0f;
    float ca = (float)collimator * DH_PI_F / 180.0f;                   // IMRTClasses.cu:35
    g.singa = sinf(ga); g.cosga = cosf(ga);
    g.sinta = sinf(ta); g.costa = cosf(ta);
    g.sinca = sinf(ca); g.cosca = cosf(ca);
    float xg = -src_dist * g.singa;                                     // CudaClasses.cu:36-49
    float yg = src_dist * g.cosga;
    g.src[0] = xg * g.costa;
    g.src[1] = yg;
    g.src[2] = -xg * g.sinta;
    return g;
}

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

dh_geom dh_photon_geom(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m)
{
    return dh_make_geom(*grid, cp->iso, cp->ga, cp->ta, cp->ca, (float)m->primary_source_distance);
}

void dh_check_photon_model(const dh_photon_model *m)
{
    DH_REQUIRE(m, "null photon model");
    DH_REQUIRE(m->n_profile > 0 && m->n_profile <= DH_MAX_PROFILE, "profile must have 1..32 points");
    DH_REQUIRE(m->n_spectral >= 0 && m->n_spectral <= DH_MAX_SPECTRAL, "spectrum must have at most 16 energies");
    DH_REQUIRE(m->profile_radius && m->profile_intensities && m->profile_softening, "null profile");
    DH_REQUIRE(m->n_spectral == 0 || (m->spectrum_attenuation_coefficients &&
               m->spectrum_primary_weights && m->spectrum_scatter_weights), "null spectrum");
}

dh_terma_args dh_make_terma_args(const dh_grid *grid, const dh_photon_cp *cp,
                                     const dh_photon_model *m, const float *mlc_rows, int32_t n_pairs)
{
    dh_check_photon_model(m);
    DH_REQUIRE(n_pairs >= 0 && n_pairs <= DH_MAX_PAIRS, "at most 128 MLC leaf pairs are supported");
    DH_REQUIRE(n_pairs == 0 || mlc_rows, "null MLC array");
    dh_terma_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = dh_photon_geom(grid, cp, m);
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

dh_ccc_args dh_make_ccc_args(const dh_grid *grid, const dh_photon_cp *cp, const dh_photon_model *m)
{
    dh_check_photon_model(m);
    dh_ccc_args a;
    std::memset(&a, 0, sizeof(a));
    a.g = dh_photon_geom(grid, cp, m);
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
            const float hx = sth * cx, hy = sth * sx;
            a.hdir[3 * (r * 12 + az) + 0] = hx;
            a.hdir[3 * (r * 12 + az) + 1] = hy;
            a.hdir[3 * (r * 12 + az) + 2] = cth;
            dh_head_to_img(a.g, hx, hy, cth, dx, dy, dz);
            a.dir[3 * (r * 12 + az) + 0] = dx;
            a.dir[3 * (r * 12 + az) + 1] = dy;
            a.dir[3 * (r * 12 + az) + 2] = dz;
        }
    }
    return a;
}

// Conservative per-layer bounds on the total sigma and on the dose factor
// P + H over every voxel the layer can reach (w in [0, 1.1 R80], source
// distance within the volume's range), from the tables themselves.
void dh_infl_layer_bounds(const dh_layer_k &L, const float4 *lut, float dist_min, float dist_max,
                         float &sig_max, float &dose_max)
{
    const float r80 = L.r80;
    float s_ms_max = -1e30f, s_ms_min = 1e30f, idd_max = 0.0f;
    for (int q = 0; q < DH_LUT_LENGTH; q++) {
        if (lut[q].x > L.cut + 1.0f && q > 0) break;
        s_ms_max = std::max(s_ms_max, lut[q].y);
        s_ms_min = std::min(s_ms_min, lut[q].y);
        idd_max = std::max(idd_max, lut[q].z);
    }
    float sa_max = -1e30f, sa_min = 1e30f;
    for (int q = 0; q <= 64; q++) {
        const float w = L.cut * q / 64.0f;
        for (int e = 0; e < 2; e++) {
            const float d = (e ? dist_max : dist_min) - w + 0.7f * r80;
            const float sa = L.A * d * d + L.B * d + L.C;
            sa_max = std::max(sa_max, sa);
            sa_min = std::min(sa_min, sa);
        }
    }
    float sh_max = 0.0f;
    for (int q = 0; q <= 256; q++) {
        const float wc = std::min(std::max(L.cut * q / 256.0f, 0.1f), r80 - 0.1f);
        const float t5 = wc - r80 - 5.0f;
        float sh = 2.85f + 0.0014f * r80 * logf(wc + 3.0f) + 0.06f * wc - 7.4e-5f * wc * wc
                 - 0.22f * r80 / (t5 * t5);
        sh_max = std::max(sh_max, sh);
    }
    const float st_max = std::max(0.0f, sa_max + s_ms_max);
    const float st_min = std::max(1e-3f, sa_min + s_ms_min);
    sig_max = 1.05f * std::sqrt(st_max * st_max + sh_max * sh_max) + 0.5f;
    // P + H <= idd / (2 pi st^2) (both normalisations are at most that)
    dose_max = 1.05f * idd_max / (2.0f * DH_PI_F * st_min * st_min);
}

// Depth-resolved table bounds of one layer: for each 1 mm cell [d_q, d_q+1]
// of the depth table within the range cut, bounds over every WET in the
// cell and every source distance in [dist_min, dist_max] of
// {sigma_t max, P max, sigma_ht max, H max} (the formulas of
// dh_pencil_factors: IDD and sigma_ms interpolate linearly, so their extremes
// are at the cell ends; sigma_air is a quadratic in its argument; the halo
// forms are sampled at the cell ends and middle and widened by 2%).
std::vector<float4> dh_infl_depth_cells(const dh_layer_k &L, const float4 *lut, float dist_min, float dist_max)
{
    std::vector<float4> out;
    const float r80 = L.r80;
    auto halo = [&](float w, float &sh, float &hw) {
        const float wc = std::min(std::max(w, 0.1f), r80 - 0.1f);
        const float t5 = wc - r80 - 5.0f, r3 = r80 + 3.0f;
        sh = 2.85f + 0.0014f * r80 * logf(wc + 3.0f) + 0.06f * wc - 7.4e-5f * wc * wc - 0.22f * r80 / (t5 * t5);
        hw = 0.052f * logf(1.13f + wc / (11.2f - 0.023f * r80)) +
             0.35f * (0.0017f * r80 * r80 - r80) / (r3 * r3 - wc * wc) - 1.61e-9f * wc * r3 * r3;
        sh = std::max(sh, 0.0f);
        hw = std::min(std::max(hw, 0.0f), 0.9f);
    };
    for (int q = 0; q + 1 < DH_LUT_LENGTH; q++) {
        // the first cell also covers the small negative WET of voxels next
        // to the upstream face (-0.05 cm + one sample), which read the first
        // table entry
        const float w0 = q == 0 ? std::min(lut[0].x, 0.0f) - 1.0f : lut[q].x;
        const float w1 = std::min(lut[q + 1].x, L.cut);
        if (q > 0 && w0 > L.cut) break;
        if (w1 < w0) continue;
        const float idd = std::max(std::max(lut[q].z, lut[q + 1].z), 0.0f);
        const float sm0 = std::min(lut[q].y, lut[q + 1].y), sm1 = std::max(lut[q].y, lut[q + 1].y);
        const float a0 = dist_min - w1 + 0.7f * r80, a1 = dist_max - w0 + 0.7f * r80;
        auto sa = [&](float d) { return (L.A * d + L.B) * d + L.C; };
        float sa0 = std::min(sa(a0), sa(a1)), sa1 = std::max(sa(a0), sa(a1));
        if (L.A != 0.0f) {
            const float dv = -L.B / (2.0f * L.A);
            if (dv > a0 && dv < a1) { sa0 = std::min(sa0, sa(dv)); sa1 = std::max(sa1, sa(dv)); }
        }
        const float st_lo = sa0 + sm0, st_hi = sa1 + sm1;
        // |sigma_t| range (P uses sigma_t^2); a range through 0 has no bound
        const float at_min = (st_lo <= 0.0f && st_hi >= 0.0f) ? 0.0f : std::min(std::fabs(st_lo), std::fabs(st_hi));
        const float at_max = 1.01f * std::max(std::fabs(st_lo), std::fabs(st_hi));
        float sh_lo = 1e30f, sh_hi = 0.0f, hw_lo = 1.0f, hw_hi = 0.0f;
        for (int e = 0; e < 3; e++) {
            float sh, hw;
            halo(e == 0 ? w0 : e == 1 ? w1 : 0.5f * (w0 + w1), sh, hw);
            sh_lo = std::min(sh_lo, sh); sh_hi = std::max(sh_hi, sh);
            hw_lo = std::min(hw_lo, hw); hw_hi = std::max(hw_hi, hw);
        }
        sh_lo *= 0.98f; sh_hi *= 1.02f;
        hw_lo = std::max(0.0f, hw_lo - 0.005f); hw_hi = std::min(0.9f, hw_hi + 0.005f);
        const float sht_min = std::sqrt(at_min * at_min + sh_lo * sh_lo);
        const float sht_max = 1.01f * std::sqrt(at_max * at_max + sh_hi * sh_hi);
        const float P = at_min > 0.0f ? 1.02f * (1.0f - hw_lo) * idd / (2.0f * DH_PI_F * at_min * at_min) : INFINITY;
        const float H = sht_min > 0.0f ? 1.02f * hw_hi * idd / (2.0f * DH_PI_F * sht_min * sht_min) : INFINITY;
        out.push_back(make_float4(at_max, P, sht_max, H));
    }
    return out;
}

// Radius beyond which every voxel of the layer, whatever its WET, has a dose
// below thr: both Gaussian terms below thr / 2 in every depth cell.
float dh_infl_depth_radius(const std::vector<float4> &cells, float thr)
{
    if (!(thr > 0.0f)) return INFINITY;
    float R = 0.0f;
    for (const float4 &c : cells) {
        if (!std::isfinite(c.y) || !std::isfinite(c.w)) return INFINITY;
        if (2.0f * c.y > thr) R = std::max(R, c.x * std::sqrt(2.0f * std::log(2.0f * c.y / thr)));
        if (2.0f * c.w > thr) R = std::max(R, c.z * std::sqrt(2.0f * std::log(2.0f * c.w / thr)));
    }
    return 1.002f * R;
}

