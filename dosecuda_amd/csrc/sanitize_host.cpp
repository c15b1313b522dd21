// sanitize_host.cpp -- CPU-only driver of the GPU-free host set-up
// (host_tables.cpp) for an AddressSanitizer + UndefinedBehaviorSanitizer
// build (dosecuda_amd.build.build_sanitize_host): spot sorting and layer
// import with ids in and out of range, TERMA arguments with closed, equal-tip,
// shuffled and maximal MLC apertures (the rectangle merge), cone arguments,
// and the influence matrix's table bounds, on seeded random inputs.  Checks
// memory and arithmetic hygiene; the results are checked by the GPU tests.
#include "dh_runtime.h"

#include <cstdio>
#include <random>
#include <vector>

int dh_set_error(int code, const char *) { return code; }   // the library's version lives in api.cpp

int main()
{
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    auto u = [&](float a, float b) { return a + (b - a) * U(rng); };
    long cases = 0;
    // proton tables
    for (int t = 0; t < 40; t++) {
        const int n_e = 1 + t % 7, stride = 400 + t % 5;
        std::vector<float> dvp(5 * n_e), dep(n_e * stride), sig(n_e * stride), idd(n_e * stride);
        for (int e = 0; e < n_e; e++) {
            dvp[5 * e] = 70.0f + e; dvp[5 * e + 1] = e == 0 ? -0.65f : 10.0f * e;
            dvp[5 * e + 2] = 1e-6f; dvp[5 * e + 3] = 1e-3f; dvp[5 * e + 4] = 3.0f;
            for (int q = 0; q < stride; q++) {
                dep[e * stride + q] = (float)q;
                sig[e * stride + q] = 0.01f * q;
                idd[e * stride + q] = (t % 9 == 0 && q == 3) ? -1.0f : 1.0f + 0.001f * q;
            }
        }
        dh_proton_lut lut{n_e, stride, dvp.data(), dep.data(), sig.data(), idd.data()};
        dh_proton_beam beam{{0, 0, 0}, 0.0, 0.0, 1340.0, 1930.0};
        const int n = t % 13 == 0 ? 0 : 1 + (int)(U(rng) * 300);
        std::vector<float> rows(4 * (size_t)n);
        for (int q = 0; q < n; q++) {
            rows[4 * q] = u(-80, 80); rows[4 * q + 1] = u(-80, 80); rows[4 * q + 2] = u(0, 2);
            rows[4 * q + 3] = (float)((int)(U(rng) * (n_e + 4)) - 2);
        }
        dh_proton_tables tb = dh_build_proton_tables(&beam, &lut, rows.data(), n);
        // influence table bounds of every layer
        for (int l = 0; l < tb.n_layers; l++) {
            float sm, dm;
            dh_infl_layer_bounds(tb.layers[l], &tb.luts[(size_t)l * DH_LUT_LENGTH], 1500.0f, 1800.0f, sm, dm);
            std::vector<float4> cells =
                dh_infl_depth_cells(tb.layers[l], &tb.luts[(size_t)l * DH_LUT_LENGTH], 1500.0f, 1800.0f);
            (void)dh_infl_depth_radius(cells, 1e-6f);
            (void)dh_infl_depth_radius(cells, 0.0f);
        }
        cases++;
    }
    // photon arguments
    std::vector<float> pr = {0, 20, 60, 120, 200, 300}, pi = {1, 1, 0.98f, 0.9f, 0.5f, 0.1f},
                       ps = {1, 1, 1, 1.1f, 1.2f, 1.3f};
    std::vector<float> mu_e = {0.05f, 0.03f, 0.02f}, pw = {0.3f, 0.4f, 0.3f}, sw = {0.2f, 0.5f, 0.3f};
    dh_photon_model m{};
    m.n_profile = 6; m.profile_radius = pr.data(); m.profile_intensities = pi.data(); m.profile_softening = ps.data();
    m.n_spectral = 3; m.spectrum_attenuation_coefficients = mu_e.data(); m.spectrum_primary_weights = pw.data();
    m.spectrum_scatter_weights = sw.data();
    m.mu_calibration = 0.008; m.primary_source_distance = 1000; m.scatter_source_distance = 850;
    m.primary_source_size = 0.6; m.scatter_source_size = 20; m.mlc_distance = 500; m.scatter_source_weight = 0.08;
    m.electron_attenuation = 0.5; m.electron_source_weight = 0.01; m.electron_fitted_dmax = 15;
    for (int r = 0; r < 6; r++) {
        m.kernel[6 * r] = 3.75f + 7.5f * r; m.kernel[6 * r + 1] = 0.01f; m.kernel[6 * r + 2] = 1.5f;
        m.kernel[6 * r + 3] = 0.002f; m.kernel[6 * r + 4] = 0.3f; m.kernel[6 * r + 5] = 1.2f + 0.4f * r;
    }
    for (int t = 0; t < 60; t++) {
        const int np = t % 11 == 0 ? 0 : (t % 7 == 0 ? DH_MAX_PAIRS : 1 + (int)(U(rng) * 100));
        std::vector<float> mlc(4 * (size_t)np);
        for (int p = 0; p < np; p++) {
            float x1 = u(-60, 20), x2 = x1 + u(-5, 50);
            if (p % 3 == 0) x2 = x1;                  // closed
            if (t % 2 && p > 0) { x1 = mlc[4 * (p - 1)]; x2 = mlc[4 * (p - 1) + 1]; }   // equal tips: a run
            mlc[4 * p] = x1; mlc[4 * p + 1] = x2;
            mlc[4 * p + 2] = -200.0f + 5.0f * p; mlc[4 * p + 3] = 5.0f;
        }
        if (t % 5 == 0)
            for (int p = np - 1; p > 0; p--) {          // shuffled pairs
                const int q = (int)(U(rng) * (p + 1));
                for (int c = 0; c < 4; c++) std::swap(mlc[4 * p + c], mlc[4 * q + c]);
            }
        dh_grid grid{20 + t % 5, 30, 25 + t % 3, u(1.5f, 3.0f), {0, 0, 0}};
        dh_photon_cp cp{{10, 20, 15}, 100.0, u(0, 360), u(-30, 30), u(-20, 20)};
        dh_terma_args ta = dh_make_terma_args(&grid, &cp, &m, mlc.data(), np);
        if (ta.n_pairs < 0 || ta.n_pairs > np) return 2;
        dh_ccc_args ca = dh_make_ccc_args(&grid, &cp, &m);
        if (ca.nsteps[0] <= 0) return 3;
        cases++;
    }
    // invalid arguments are refused with dh_error, not undefined behaviour
    int refused = 0;
    try {
        dh_proton_lut bad{2, 100, nullptr, nullptr, nullptr, nullptr};
        dh_proton_beam beam{{0, 0, 0}, 0.0, 0.0, 1340.0, 1930.0};
        (void)dh_build_proton_tables(&beam, &bad, nullptr, 0);
    } catch (const dh_error &) { refused++; }
    try {
        std::vector<float> mlc(4 * (DH_MAX_PAIRS + 1), 1.0f);
        dh_grid grid{4, 4, 4, 2.0f, {0, 0, 0}};
        dh_photon_cp cp{{0, 0, 0}, 100.0, 0.0, 0.0, 0.0};
        (void)dh_make_terma_args(&grid, &cp, &m, mlc.data(), DH_MAX_PAIRS + 1);
    } catch (const dh_error &) { refused++; }
    if (refused != 2) return 4;
    std::printf("sanitize host: ok (%ld cases, %d refusals)\n", cases, refused);
    return 0;
}
