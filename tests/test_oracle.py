"""Pins of the FP64 oracle (CPU only).

Each check is chosen so that a plausible transcription slip (axis order, sign
of an angle, a dropped factor, a wrong border value, an off-by-one step) fails
it, using closed forms, invariants, library routines and an independent
NumPy restatement (oracle/mini.py) on small non-cubic grids.
"""
import math

import numpy as np
import pytest

from oracle import mini
from oracle import oracle as O
from tests.conftest import rel_err

F = np.float32


# ---------------------------------------------------------------------------
# texture sampling (TextureClasses.cu:14-41)
# ---------------------------------------------------------------------------
def test_sample_exact_at_texel_centres():
    rng = np.random.default_rng(0)
    v = rng.random((4, 5, 6)).astype(F)
    for (k, j, i) in [(0, 0, 0), (3, 4, 5), (2, 1, 3)]:
        assert O.sample(v, (i + 0.5, j + 0.5, k + 0.5), 7.0) == pytest.approx(float(v[k, j, i]), abs=0)


def test_sample_border_blend_half_texel():
    v = np.full((3, 3, 3), 2.0, F)
    b = 0.25
    # t in [0, 0.5): weight (0.5 + t) on voxel 0, (0.5 - t) on the border
    for t in (0.0, 0.1, 0.3, 0.49):
        got = O.sample(v, (t, 1.5, 1.5), b)
        assert got == pytest.approx((0.5 + t) * 2.0 + (0.5 - t) * b, rel=1e-12)
    # far outside: border
    assert O.sample(v, (-5.0, 1.5, 1.5), b) == pytest.approx(b)
    assert O.sample(v, (1.5, 1.5, 9.0), b) == pytest.approx(b)


def test_sample_matches_independent_trilinear():
    rng = np.random.default_rng(1)
    v = rng.random((5, 6, 7)).astype(F)
    pts = rng.uniform(-1.5, 8.5, size=(200, 3))
    ref = mini.sample(v, pts[:, 0], pts[:, 1], pts[:, 2], 0.3)
    got = np.array([O.sample(v, p, 0.3) for p in pts])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------------------
# geometry (CudaClasses.cu:17-51)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("gantry, expect", [(0.0, (0, -1, 0)), (90.0, (1, 0, 0)),
                                            (180.0, (0, 1, 0)), (270.0, (-1, 0, 0))])
def test_source_position_cardinal(gantry, expect):
    g = O.geom((4, 4, 4), 1.0, np.zeros(3), np.zeros(3), gantry, 0.0, 0.0, 1000.0)
    np.testing.assert_allclose(np.array(g.src) / 1000.0, expect, atol=1e-6)


def test_source_couch_rotates_about_y():
    g = O.geom((4, 4, 4), 1.0, np.zeros(3), np.zeros(3), 90.0, 90.0, 0.0, 1000.0)
    # gantry 90 puts the source on +x; couch 90 turns x into -z
    np.testing.assert_allclose(np.array(g.src) / 1000.0, (0, 0, -1), atol=1e-6)


# ---------------------------------------------------------------------------
# ray trace (CudaClasses.cu:180-228)
# ---------------------------------------------------------------------------
def _closed_form_column_wet(rho_line, s, j):
    """WET of voxel j of a column along -y: samples at L = 0, 1, 2 ... mm,
    texel coordinate u = j - L/s, stop once t = u + 0.5 < 0; linear
    interpolation along y with border 0 below voxel 0."""
    acc = 0.0
    L = 0
    while True:
        u = j - L / s
        if u + 0.5 < 0:
            break
        i0 = math.floor(u)
        f = u - i0
        a = rho_line[i0] if i0 >= 0 else 0.0
        b = rho_line[i0 + 1] if i0 + 1 < len(rho_line) else 0.0
        acc += max((1 - f) * a + f * b, 0.0)
        L += 1
    return -0.05 + 0.1 * acc


@pytest.mark.parametrize("s", [2.0, 2.5, 3.0])
def test_raytrace_slab_closed_form(s):
    # (nz, ny, nx) non-cubic; density varies along y only; iso on the column
    nz, ny, nx = 5, 13, 7
    prof = np.array([1.0, 1.0, 1.6, 1.6, 0.3, 0.3, 1.0, 1.2, 1.2, 1.0, 0.9, 1.0, 1.0], F)
    rho = np.broadcast_to(prof[None, :, None], (nz, ny, nx)).copy()
    iso = np.array([3 * s, 6 * s, 2 * s], F)
    g = O.geom(rho.shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, 1635.0)
    wet, ns = O.raytrace(g, rho, 0.0)
    for j in range(ny):
        exp = _closed_form_column_wet(prof.astype(np.float64), s, j)
        assert wet[2, j, 3] == pytest.approx(exp, abs=2e-6), j
        assert ns[2, j, 3] == math.floor(s * (j + 0.5)) + 1


def test_raytrace_gantry90_runs_along_plus_x():
    nz, ny, nx = 3, 4, 9
    rho = np.ones((nz, ny, nx), F)
    s = 2.0
    iso = np.array([4 * s, 1.5 * s, 1 * s], F)
    g = O.geom(rho.shape, s, np.zeros(3), iso, 90.0, 0.0, 0.0, 1635.0)
    wet, ns = O.raytrace(g, rho, 0.0)
    # rays go to +x: voxel nx-1 has the shortest ray
    row = wet[1, 1, :]
    assert np.all(np.diff(row) < 0)


def test_raytrace_matches_mini_oblique():
    rng = np.random.default_rng(3)
    shape = (12, 16, 20)
    rho = rng.uniform(0.2, 1.8, shape).astype(F)
    iso = np.array([17.0, 13.0, 11.0], F)
    origin = np.array([-2.0, 1.0, 0.5], F)
    for (ga, ta) in [(33.0, 0.0), (200.0, 15.0), (271.0, 350.0)]:
        g = O.geom(shape, 2.0, origin, iso, ga, ta, 0.0, 1635.0)
        w, _ = O.raytrace(g, rho, 0.0012)
        mg = mini.G(shape, 2.0, origin, iso, ga, ta, 0.0, 1635.0)
        mg.set_trig(g.singa, g.cosga, g.sinta, g.costa, g.sinca, g.cosca)
        mg.src = np.array(g.src, F)
        wm = mini.raytrace(mg, rho, 0.0012)
        np.testing.assert_allclose(w, wm, rtol=0, atol=1e-9)


# ---------------------------------------------------------------------------
# smoothing (IMPTClasses.cu:139-202)
# ---------------------------------------------------------------------------
def test_smooth_constant_and_shallow():
    shape = (9, 10, 11)
    g = O.geom(shape, 2.0, np.zeros(3), np.array([10.0, 9.0, 8.0], F), 0.0, 0.0, 0.0, 1635.0)
    c = np.full(shape, 2.0, F)
    np.testing.assert_allclose(O.smooth(g, c), 2.0, rtol=1e-15)
    # WET < 0.1 cm: no neighbour radius passes r < 10 * WET
    c = np.full(shape, 0.09, F)
    c[4, 5, 5] = 0.05
    out = O.smooth(g, c)
    np.testing.assert_allclose(out, c.astype(np.float64), rtol=0, atol=0)


def test_smooth_counts_55_samples():
    # one perturbed voxel inside a deep homogeneous WET: the voxel itself sees
    # 1 + 6 directions x 9 radii = 55 samples
    shape = (41, 41, 41)
    s = 1.0
    iso = np.array([20.0, 20.0, 20.0], F)
    g = O.geom(shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, 1635.0)
    c = np.full(shape, 3.0, F)
    c[20, 20, 20] = 3.0 + 55.0
    out = O.smooth(g, c)
    assert out[20, 20, 20] == pytest.approx(3.0 + 1.0, rel=1e-12)


def test_smooth_matches_mini():
    rng = np.random.default_rng(4)
    shape = (12, 16, 20)
    wet = rng.uniform(0.0, 1.5, shape).astype(F)
    iso = np.array([17.0, 13.0, 11.0], F)
    g = O.geom(shape, 2.0, np.zeros(3), iso, 47.0, 10.0, 0.0, 1635.0)
    mg = mini.G(shape, 2.0, np.zeros(3), iso, 47.0, 10.0, 0.0, 1635.0)
    mg.set_trig(g.singa, g.cosga, g.sinta, g.costa, g.sinca, g.cosca)
    np.testing.assert_allclose(O.smooth(g, wet), mini.smooth(mg, wet), rtol=1e-12)


# ---------------------------------------------------------------------------
# pencil beam (IMPTClasses.cu:69-133,204-271)
# ---------------------------------------------------------------------------
def _synthetic_model(n_e=3, r80=150.0, A=0.0, B=0.0, C=5.0, idd=lambda d: 1.0 + 0.01 * d,
                     sig=lambda d: 0.0 * d):
    depth = np.arange(400, dtype=F)
    dvp = np.array([[100.0 + e, r80, A, B, C] for e in range(n_e)], F)
    lut_d = np.tile(depth, (n_e, 1))
    lut_i = np.tile(np.array([idd(d) for d in depth], F), (n_e, 1))
    lut_s = np.tile(np.array([sig(d) for d in depth], F), (n_e, 1))
    return dvp, lut_d, lut_s, lut_i


def test_pencil_lateral_integral_is_idd():
    # plane of voxels at fixed depth, 1 mm spacing, one spot on the axis:
    # sum(dose) * dx^2 = MU * IDD(w)  (normalised Gaussians, (1-w) + w = 1)
    nz, ny, nx = 121, 1, 121
    s = 1.0
    iso = np.array([60.0, 0.0, 60.0], F)
    wet = np.full((nz, ny, nx), 5.0, F)              # 50 mm
    dvp, d, sg, idd = _synthetic_model(idd=lambda x: 3.0 + 0.02 * x)
    spots = np.array([[0.0, 0.0, 2.0, 1]], F)
    g = O.geom(wet.shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, (1340.0 + 1930.0) / 2)
    dose, pairs = O.pencil(g, 1340.0, 1930.0, wet, spots, dvp, d, sg, idd)
    assert pairs == nz * ny * nx
    assert dose.sum() * s * s == pytest.approx(2.0 * (3.0 + 0.02 * 50.0), rel=2e-4)


def test_pencil_pure_gaussian_and_cax_distance():
    # R80 = 5 mm forces the halo weight to 0; sigma = C exactly.  Dose is then
    # MU*IDD/(2 pi C^2) exp(-d^2 / 2C^2) with d the point-line distance to the
    # divergent spot axis, computed here with a cross product.
    nz, ny, nx = 9, 3, 11
    s = 2.0
    iso = np.array([10.0, 2.0, 8.0], F)
    wet = np.full((nz, ny, nx), 0.3, F)              # 3 mm < 1.1 * R80
    C = 4.0
    dvp, d, sg, idd = _synthetic_model(r80=5.0, C=C, idd=lambda x: 2.0)
    x0, y0 = 3.0, -2.0
    spots = np.array([[x0, y0, 1.5, 2]], F)
    vx, vy = 1340.0, 1930.0
    g = O.geom(wet.shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, (vx + vy) / 2)
    dose, _ = O.pencil(g, vx, vy, wet, spots, dvp, d, sg, idd)
    # BEV coordinates at gantry 0 (adjusted 180): head = (x, z, -y)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    px, py, pz = i * s - iso[0], j * s - iso[1], k * s - iso[2]
    h = np.stack([px, pz, -py], -1).astype(np.float64)
    tau = np.array([x0 / vx, y0 / vy, -1.0])
    r = h - np.array([x0, y0, 0.0])
    d2 = (np.cross(r, tau) ** 2).sum(-1) / (tau @ tau)
    exp = 1.5 * 2.0 / (2 * np.pi * C * C) * np.exp(-d2 / (2 * C * C))
    np.testing.assert_allclose(dose, exp, rtol=1e-6)


def test_pencil_range_cut_and_layers(proton_model):
    m = proton_model
    shape = (6, 7, 8)
    wet = np.linspace(0.0, 30.0, np.prod(shape)).reshape(shape).astype(F)
    g = O.geom(shape, 2.5, np.zeros(3), np.array([8.0, 7.0, 6.0], F), 0.0, 0.0, 0.0,
               (m.VSADX + m.VSADY) / 2)
    spots = np.array([[0, 0, 1, 40], [5, 5, 1, 10], [1, 1, 1, 40]], F)
    dose, pairs = O.pencil(g, m.VSADX, m.VSADY, wet, spots, m.divergence_params, m.lut_depths,
                           m.lut_sigmas, m.lut_idds)
    w = wet * F(10.0)
    exp_pairs = 2 * np.count_nonzero(~(w > F(1.1) * m.divergence_params[40, 1])) + \
        np.count_nonzero(~(w > F(1.1) * m.divergence_params[10, 1]))
    assert pairs == exp_pairs
    assert np.all(dose[w > F(1.1) * max(m.divergence_params[40, 1], m.divergence_params[10, 1])] == 0)


def test_pencil_layer_import_quirks():
    # importLayers counts contiguous runs from energy id 0: a negative id sorts
    # first and empties every layer; ids >= n_energies are dropped.
    dvp, d, sg, idd = _synthetic_model(n_e=3)
    _, layers, nl = O.make_layers(np.array([[0, 0, 1, 1], [0, 0, 1, -1]], F), 3, dvp)
    assert nl == 0
    _, layers, nl = O.make_layers(np.array([[0, 0, 1, 2], [0, 0, 1, 7], [0, 0, 1, 0]], F), 3, dvp)
    assert nl == 2 and layers[0].id == 0 and layers[1].id == 2 and layers[1].n == 1


def test_pencil_linear_in_mu_and_permutation(proton_model):
    m = proton_model
    rng = np.random.default_rng(5)
    shape = (10, 12, 14)
    wet = rng.uniform(0.0, 12.0, shape).astype(F)
    g = O.geom(shape, 2.5, np.zeros(3), np.array([16.0, 14.0, 12.0], F), 30.0, 0.0, 0.0,
               (m.VSADX + m.VSADY) / 2)
    spots = np.column_stack([rng.uniform(-20, 20, 30), rng.uniform(-20, 20, 30),
                             rng.uniform(0.5, 2, 30), rng.integers(20, 25, 30)]).astype(F)
    args = (m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
    d1, _ = O.pencil(g, m.VSADX, m.VSADY, wet, spots, *args)
    s2 = spots.copy()
    s2[:, 2] *= 2
    d2, _ = O.pencil(g, m.VSADX, m.VSADY, wet, s2, *args)
    np.testing.assert_allclose(d2, 2 * d1, rtol=1e-12)
    d3, _ = O.pencil(g, m.VSADX, m.VSADY, wet, spots[rng.permutation(30)], *args)
    np.testing.assert_allclose(d3, d1, rtol=1e-12)


def test_pencil_matches_mini(proton_model):
    m = proton_model
    rng = np.random.default_rng(6)
    shape = (12, 16, 20)
    wet = rng.uniform(0.0, 14.0, shape).astype(F)
    iso = np.array([20.0, 15.0, 12.0], F)
    spots = np.column_stack([rng.uniform(-15, 15, 12), rng.uniform(-15, 15, 12),
                             rng.uniform(0.5, 2, 12), rng.integers(30, 34, 12)]).astype(F)
    spots = spots[np.argsort(spots[:, 3], kind="stable")]
    g = O.geom(shape, 2.5, np.zeros(3), iso, 120.0, 5.0, 0.0, (m.VSADX + m.VSADY) / 2)
    mg = mini.G(shape, 2.5, np.zeros(3), iso, 120.0, 5.0, 0.0, (m.VSADX + m.VSADY) / 2)
    mg.set_trig(g.singa, g.cosga, g.sinta, g.costa, g.sinca, g.cosca)
    mg.src = np.array(g.src, F)
    args = (m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
    d1, _ = O.pencil(g, m.VSADX, m.VSADY, wet, spots, *args)
    d2 = mini.pencil(mg, m.VSADX, m.VSADY, wet, spots, *args)
    assert rel_err(d1, d2)[0] < 1e-9


# ---------------------------------------------------------------------------
# TERMA (IMRTClasses.cu:120-171,209-250)
# ---------------------------------------------------------------------------
class _PM:
    def __init__(self, base, **kw):
        for k in ("mu_calibration", "primary_source_distance", "scatter_source_distance",
                  "primary_source_size", "scatter_source_size", "mlc_distance",
                  "scatter_source_weight", "electron_attenuation", "electron_source_weight",
                  "electron_fitted_dmax", "profile_radius", "profile_intensities",
                  "profile_softening", "spectrum_attenuation_coefficients",
                  "spectrum_primary_weights", "spectrum_scatter_weights", "kernel"):
            setattr(self, k, getattr(base, k))
        for k, v in kw.items():
            setattr(self, k, v)


def _open_mlc(n=60, width=10.0, half=200.0):
    off = (np.arange(n) - (n - 1) / 2) * width
    return np.column_stack([np.full(n, -half), np.full(n, half), off, np.full(n, width)]).astype(F)


def test_terma_open_field_transmission_and_closed(photon_model):
    shape = (3, 5, 3)
    iso = np.array([2.0, 4.0, 2.0], F)
    g = O.geom(shape, 2.0, np.zeros(3), iso, 0.0, 0.0, 0.0, 1000.0)
    pm = _PM(photon_model, scatter_source_weight=0.0,
             spectrum_attenuation_coefficients=np.array([0.05], F),
             spectrum_primary_weights=np.array([0.0], F), spectrum_scatter_weights=np.array([1.0], F),
             profile_intensities=np.ones(20, F), profile_softening=np.ones(20, F))
    wet = np.zeros(shape, F)
    t, e = O.terma(g, pm, _open_mlc(), wet)
    # open field 400 x 600 mm >> sigma: Tp = Ts = 1, tr = 1, el(W=0) = 1
    # TERMA = sw * SSD_s / (SSD_s - h_z);  E = esw * 0.7 * Tp
    hz = -(np.arange(5) * 2.0 - 4.0)              # gantry 0: h_z = -y
    exp = 850.0 / (850.0 - hz)
    np.testing.assert_allclose(t[1, :, 1], exp, rtol=1e-7)   # erf tails of the 30 mm scatter source
    np.testing.assert_allclose(e, 2.0 * 0.7, rtol=1e-7)
    # Beer-Lambert with unit softening on the axis
    wet2 = np.full(shape, 3.0, F)
    t2, e2 = O.terma(g, pm, _open_mlc(), wet2)
    np.testing.assert_allclose(t2 / t, math.exp(-float(F(0.05)) * 3.0), rtol=1e-9)
    # electron contamination vanishes beyond the fitted d_max
    wet3 = np.full(shape, 1.6, F)
    assert np.all(O.terma(g, pm, _open_mlc(), wet3)[1] == 0.0)
    # closed leaves: no transmission, no TERMA
    closed = _open_mlc()
    closed[:, 1] = closed[:, 0]
    t0, e0 = O.terma(g, pm, closed, wet)
    assert np.all(t0 == 0.0) and np.all(e0 == 0.0)


def test_terma_field_edge_is_erf(photon_model):
    # one open pair, voxel plane at iso depth: Tp = 1/4 (erf(X2)-erf(X1))(erf(Y2)-erf(Y1))
    # with edges projected through the MLC plane to the source plane.
    shape = (1, 1, 41)
    s = 1.0
    iso = np.array([20.0, 0.0, 0.0], F)
    g = O.geom(shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, 1000.0)
    pm = _PM(photon_model, scatter_source_weight=0.0, spectrum_attenuation_coefficients=np.array([0.0], F),
             spectrum_primary_weights=np.array([1.0], F), spectrum_scatter_weights=np.array([0.0], F),
             scatter_source_size=1e6,
             profile_intensities=np.ones(20, F), profile_softening=np.ones(20, F))
    mlc = np.array([[-10.0, 15.0, 0.0, 40.0]], F)
    t, e = O.terma(g, pm, mlc, np.zeros(shape, F))
    from scipy.special import erf
    x = np.arange(41) * s - 20.0
    hx = x                                      # gantry 0 (adjusted 180): head = (x, z, -y)
    sad, dm, sig = 1000.0, 491.0, 1.0
    m = (sad - dm) / sad
    dv = sad / dm                               # (Z - h_z)/(d_mlc - h_z) at h_z = 0
    X1 = (-10.0 * m - hx) * dv + hx
    X2 = (15.0 * m - hx) * dv + hx
    Y1 = (-20.0 * m) * dv
    Y2 = (20.0 * m) * dv
    inv = 1 / (sig * math.sqrt(2))
    Tp = 0.25 * (erf(X2 * inv) - erf(X1 * inv)) * (erf(Y2 * inv) - erf(Y1 * inv))
    # scatter source of huge size: Ts ~ 0, tr ~ 0; TERMA = Tp * 1 * (SAD/SAD)^2
    np.testing.assert_allclose(t[0, 0, :], Tp, rtol=1e-6, atol=1e-9)


def _mini_model(pm, mlc):
    return dict(sad=float(F(pm.primary_source_distance)), ssd_s=float(F(pm.scatter_source_distance)),
                d_mlc=float(F(pm.mlc_distance)), src_p=float(F(pm.primary_source_size)),
                src_s=float(F(pm.scatter_source_size)), ssw=float(F(pm.scatter_source_weight)),
                el_att=float(F(pm.electron_attenuation)), el_dmax=float(F(pm.electron_fitted_dmax)),
                esw=float(F(pm.electron_source_weight)), mu_cal=float(F(pm.mu_calibration)),
                prof_r=pm.profile_radius.astype(np.float64), prof_i=pm.profile_intensities.astype(np.float64),
                prof_s=pm.profile_softening.astype(np.float64),
                mu_e=pm.spectrum_attenuation_coefficients.astype(np.float64),
                pw=pm.spectrum_primary_weights.astype(np.float64),
                sw=pm.spectrum_scatter_weights.astype(np.float64),
                kernel=np.asarray(pm.kernel, np.float64))


def test_terma_matches_mini(photon_model):
    from dosecuda_amd.beam_models import square_field_mlc
    rng = np.random.default_rng(7)
    shape = (12, 16, 20)
    wet = rng.uniform(0.0, 20.0, shape).astype(F)
    iso = np.array([20.0, 15.0, 12.0], F)
    mlc = square_field_mlc(photon_model, 3.0, 2.0)
    g = O.geom(shape, 2.5, np.zeros(3), iso, 40.0, 10.0, 25.0, 1000.0)
    mg = mini.G(shape, 2.5, np.zeros(3), iso, 40.0, 10.0, 25.0, 1000.0)
    mg.set_trig(g.singa, g.cosga, g.sinta, g.costa, g.sinca, g.cosca)
    t, e = O.terma(g, photon_model, mlc, wet)
    tm, em = mini.terma(mg, _mini_model(photon_model, mlc), mlc, wet)
    assert rel_err(t, tm)[0] < 1e-9
    assert rel_err(e, em)[0] < 1e-9


# ---------------------------------------------------------------------------
# CCCS (IMRTClasses.cu:252-350)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("s, n", [(2.5, 54), (2.0, 61), (3.0, 43)])
def test_cone_trip_counts(photon_model, s, n):
    L = photon_model.kernel[:, 5]
    assert sum(O.cone_steps(l, s) for l in L) == n


def test_ccc_geometric_series(photon_model):
    # constant TERMA T and density rho around the voxel, every ray inside:
    # Rp_N = K (1 - e^N), K = T sin(th) A / a^2, e = exp(-a rho sp)
    # Rs_N = T sin(th) B / b^2 (1 - (1 - b rho sp)^N)
    s = 2.5
    n = 33
    shape = (n, n, n)
    T, rho = 0.7, 1.1
    terma = np.full(shape, T, F)
    dens = np.full(shape, rho, F)
    el = np.full(shape, 0.05, F)
    iso = np.array([16 * s, 16 * s, 16 * s], F)
    g = O.geom(shape, s, np.zeros(3), iso, 30.0, 0.0, 10.0, 1000.0)
    mask = np.zeros(shape, np.uint8)
    mask[16, 16, 16] = 1
    dose, _ = O.ccc(g, photon_model, 100.0, terma, dens, el, mask)
    T, rho = float(F(T)), float(F(rho))
    K = np.asarray(photon_model.kernel, np.float64)
    sp = float(F(s) / F(10.0))
    tot = 0.0
    for r in range(6):
        th = float(F(K[r, 0]) * F(3.141592654) / F(180.0))
        A, a, B, b = K[r, 1:5]
        N = O.cone_steps(K[r, 5], s)
        e = math.exp(-a * rho * sp)
        Rp = T * math.sin(th) * A / a ** 2 * (1 - e ** N)
        Rs = T * math.sin(th) * B / b ** 2 * (1 - (1 - b * rho * sp) ** N)
        tot += 12 * (a * Rp + b * Rs)
    tot += float(F(0.05))
    exp = float(F(photon_model.mu_calibration)) * tot * 100.0
    assert dose[16, 16, 16] == pytest.approx(exp, rel=1e-9)


def test_ccc_early_out_and_matches_mini(photon_model):
    rng = np.random.default_rng(8)
    shape = (10, 12, 14)
    terma = rng.uniform(0.0, 1.0, shape).astype(F)
    terma[terma < 0.3] = 0.005                       # below the 0.01 early-out
    dens = rng.uniform(0.2, 1.6, shape).astype(F)
    el = rng.uniform(0.0, 0.1, shape).astype(F)
    iso = np.array([16.0, 14.0, 12.0], F)
    g = O.geom(shape, 2.5, np.zeros(3), iso, 70.0, 20.0, 15.0, 1000.0)
    dose, _ = O.ccc(g, photon_model, 50.0, terma, dens, el)
    assert np.all(dose[terma <= 0.01] == 0.0)
    mg = mini.G(shape, 2.5, np.zeros(3), iso, 70.0, 20.0, 15.0, 1000.0)
    mg.set_trig(g.singa, g.cosga, g.sinta, g.costa, g.sinca, g.cosca)
    mg.src = np.array(g.src, F)
    dm = mini.ccc(mg, _mini_model(photon_model, None), 50.0, terma, dens, el)
    assert rel_err(dose, dm)[0] < 1e-9
