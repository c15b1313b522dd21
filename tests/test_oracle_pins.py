"""Pins of the oracle's geometry and pencil-beam composition (CPU only).

These pin the parts of oracle/dose_oracle.c that the engine shares a reading
with, by properties that do not come from the same transcription:

* the image <-> head rotation chain (CudaClasses.cu:65-126, with the
  collimator step of IMRTClasses.cu:41-118): it is a proper rotation, head
  then image is the identity, the source lands on the head +z axis at the
  source distance, the collimator turns head x-y about the central axis, and
  the whole chain equals a product of library rotation matrices
  (scipy.spatial.transform);
* kernelTilt (IMRTClasses.cu:173-203): a proper rotation that maps the
  iso -> source unit vector onto the voxel -> source unit vector, fixes the
  axis a x u (so it is the minimal rotation), is the identity on the central
  axis, and equals scipy's rotation-vector construction;
* the cone directions of cccKernel (IMRTClasses.cu:282-306): at zenith theta
  every tilted direction makes the angle theta with the voxel -> source
  vector, and opposite azimuths are mirror images about it;
* sigmaAir and its argument (IMPTClasses.cu:94-102,236-237): the Gaussian
  width read back from the dose on and off the spot axis equals
  A d^2 + B d + C + sigma_ms at d = |S - p| - w + 0.7 R80, with S at the
  closed-form source position;
* the halo composition (IMPTClasses.cu:104-121,239-244): the tail width read
  back from the dose equals hypot(sigma_t, sigma_halo), the tail weight the
  halo weight, and the axis dose (1 - w) P + w H.

The empirical coefficient values of sigmaAir and nuclearHalo themselves stay
unpinned (no reference run is possible); the tests use them only through
the composition they feed.

Each ``pin_*`` function raises AssertionError when its property fails;
tests/test_oracle_mutations.py checks that every planted one-token mistake in
a copy of dose_oracle.c fails at least one of them.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import oracle as O

F = np.float32
ANGLES = [(0.0, 0.0, 0.0), (30.0, 0.0, 0.0), (0.0, 25.0, 0.0), (0.0, 0.0, 40.0),
          (51.43, 10.0, 20.0), (120.0, -35.0, 95.0), (233.0, 80.0, -60.0),
          (308.57, 355.0, 180.0), (90.0, 90.0, 90.0), (271.0, 275.0, 10.0)]
D_SRC = 1000.0


def _geom(ga, ta, ca, D=D_SRC):
    return O.geom((8, 8, 8), 2.5, np.zeros(3), np.zeros(3), ga, ta, ca, D)


def _rad(deg):
    """The float32 angle the set-up forms (dosemodule.cu:152, CudaClasses.cu:27-28)."""
    return float(F(deg) * F(3.141592654) / F(180.0))


def _independent_matrix(ga, ta, ca):
    """Image -> head as a product of library rotations: couch about image y by
    -couch, gantry about z by -(gantry + 180), collimator about y by
    -collimator, then the axis relabelling (x, y, z) -> (-x, z, y)."""
    g = _rad(float(np.fmod(F(ga + 180.0), F(360.0))))
    R = (Rotation.from_euler("y", -_rad(ca)) * Rotation.from_euler("z", -g) *
         Rotation.from_euler("y", -_rad(ta))).as_matrix()
    S = np.array([[-1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.0, 1.0, 0.0]])
    return S @ R


def _points(rng, n=12):
    return rng.uniform(-200.0, 200.0, size=(n, 3))


# ---------------------------------------------------------------------------
# rotation chain
# ---------------------------------------------------------------------------
def pin_rotation_round_trip_and_norm():
    rng = np.random.default_rng(11)
    for ga, ta, ca in ANGLES:
        g = _geom(ga, ta, ca)
        for fp32 in (False, True):
            for p in _points(rng):
                h = O.img_to_head(g, p, fp32)
                q = O.head_to_img(g, h, fp32)
                scale = np.linalg.norm(p)
                assert np.abs(q - p).max() <= 2e-6 * scale, (ga, ta, ca, fp32, p, q)
                assert abs(np.linalg.norm(h) - scale) <= 2e-6 * scale, (ga, ta, ca, fp32)


def pin_source_on_head_axis():
    # the source (CudaClasses.cu:17-51) sits on the head +z axis at distance D
    for ga, ta, ca in ANGLES:
        g = _geom(ga, ta, ca)
        for fp32 in (False, True):
            h = O.img_to_head(g, np.array(g.src, np.float64), fp32)
            np.testing.assert_allclose(h, [0.0, 0.0, D_SRC], atol=2e-6 * D_SRC,
                                       err_msg=f"{ga} {ta} {ca} fp32={fp32}")


def pin_rotation_equals_library_matrices():
    rng = np.random.default_rng(12)
    for ga, ta, ca in ANGLES:
        g = _geom(ga, ta, ca)
        M = _independent_matrix(ga, ta, ca)
        assert abs(np.linalg.det(M) - 1.0) < 1e-12
        for fp32 in (False, True):
            for p in _points(rng, 6):
                scale = np.linalg.norm(p)
                np.testing.assert_allclose(O.img_to_head(g, p, fp32), M @ p, atol=2e-6 * scale,
                                           err_msg=f"{ga} {ta} {ca} fp32={fp32}")
                np.testing.assert_allclose(O.head_to_img(g, p, fp32), M.T @ p,
                                           atol=2e-6 * scale,
                                           err_msg=f"{ga} {ta} {ca} fp32={fp32}")


def pin_collimator_turns_about_central_axis():
    # head(collimator c) = Rz(-c) head(collimator 0): head z and central-axis
    # points are unchanged, the head x-y plane turns by -c
    rng = np.random.default_rng(13)
    for ga, ta, _ in ANGLES:
        g0 = _geom(ga, ta, 0.0)
        for ca in (15.0, 90.0, 200.0):
            g1 = _geom(ga, ta, ca)
            c = _rad(ca)
            Rz = np.array([[math.cos(c), math.sin(c), 0.0], [-math.sin(c), math.cos(c), 0.0],
                           [0.0, 0.0, 1.0]])
            for fp32 in (False, True):
                for p in _points(rng, 6):
                    scale = np.linalg.norm(p)
                    np.testing.assert_allclose(O.img_to_head(g1, p, fp32),
                                               Rz @ O.img_to_head(g0, p, fp32), atol=2e-6 * scale)
                cax = 0.37 * np.array(g1.src, np.float64)
                np.testing.assert_allclose(O.img_to_head(g1, cax, fp32),
                                           [0.0, 0.0, 0.37 * D_SRC], atol=2e-6 * D_SRC)


# ---------------------------------------------------------------------------
# kernel tilt
# ---------------------------------------------------------------------------
def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _tilt_matrix(g, p):
    return np.column_stack([O.kernel_tilt(g, p, e) for e in np.eye(3)])


def pin_tilt_maps_axis_onto_voxel_direction():
    rng = np.random.default_rng(14)
    for ga, ta, ca in ANGLES:
        g = _geom(ga, ta, ca)
        a = _unit(g.src)
        for p in _points(rng, 8):
            u = _unit(np.array(g.src, np.float64) - p)
            np.testing.assert_allclose(O.kernel_tilt(g, p, a), u, atol=1e-12)
            R = _tilt_matrix(g, p)
            np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-12)
            assert abs(np.linalg.det(R) - 1.0) < 1e-12
            k = np.cross(a, u)
            np.testing.assert_allclose(R @ k, k, atol=1e-12)   # minimal rotation: axis fixed
            ang = math.acos(min(1.0, max(-1.0, float(a @ u))))
            Rs = Rotation.from_rotvec(ang * _unit(k)).as_matrix()
            np.testing.assert_allclose(R, Rs, atol=1e-10)


def pin_tilt_identity_on_central_axis():
    rng = np.random.default_rng(15)
    for ga, ta, ca in ANGLES:
        g = _geom(ga, ta, ca)
        for lam in (-0.3, 0.0, 0.2, 0.6):
            p = lam * np.array(g.src, np.float64)
            for d in rng.normal(size=(4, 3)):
                np.testing.assert_allclose(O.kernel_tilt(g, p, d), d, atol=1e-12)


# ---------------------------------------------------------------------------
# cone directions of the superposition
# ---------------------------------------------------------------------------
def pin_cone_directions(kernel):
    rng = np.random.default_rng(16)
    K = np.asarray(kernel, np.float64)
    for ga, ta, ca in ANGLES[::2]:
        g = _geom(ga, ta, ca)
        M = _independent_matrix(ga, ta, ca)
        for p in _points(rng, 3):
            u = _unit(np.array(g.src, np.float64) - p)
            a = _unit(g.src)
            k = np.cross(a, u)
            Rt = Rotation.from_rotvec(math.acos(min(1.0, float(a @ u))) * _unit(k)).as_matrix()
            for r in range(6):
                th = _rad(K[r, 0])
                for az in range(12):
                    d = O.cone_dir(g, kernel, r, az, p)
                    assert abs(np.linalg.norm(d) - 1.0) < 1e-7
                    assert abs(float(d @ u) - math.cos(th)) < 2e-7, (r, az)
                    assert abs(np.linalg.norm(np.cross(d, u)) - math.sin(th)) < 2e-7, (r, az)
                    phi = float(F(az) * F(3.141592654) / F(6.0))
                    dh = np.array([math.sin(th) * math.cos(phi), math.sin(th) * math.sin(phi),
                                   math.cos(th)])
                    np.testing.assert_allclose(d, Rt @ (M.T @ dh), atol=2e-6)
                for az in range(6):
                    s = O.cone_dir(g, kernel, r, az, p) + O.cone_dir(g, kernel, r, az + 6, p)
                    np.testing.assert_allclose(s, 2.0 * math.cos(th) * u, atol=2e-6)


# ---------------------------------------------------------------------------
# sigmaAir argument and halo composition, read back from the dose
# ---------------------------------------------------------------------------
VSADX, VSADY = 1340.0, 1930.0


def _model(r80, A, B, C, idd, sig_ms, n_e=3):
    depth = np.arange(400, dtype=F)
    dvp = np.array([[100.0 + e, r80, A, B, C] for e in range(n_e)], F)
    lut_d = np.tile(depth, (n_e, 1))
    lut_i = np.full((n_e, 400), idd, F)
    lut_s = np.full((n_e, 400), sig_ms, F)
    return dvp, lut_d, lut_s, lut_i


def _plane_dose(wet_cm, r80, A, B, C, idd, sig_ms, mu=1.5, nx=61, nz=1, s=1.0, spot=(0.0, 0.0),
                iso_z=0.0):
    """Dose of one spot on an (nz, 1, nx) plane of voxels at the isocentre
    depth (gantry 0: source at (0, -D, 0), beam along +y, head = (x, z, -y)).
    Returns the dose [nz, nx], image x and z of the voxels and |S - p|."""
    iso = np.array([0.0, 0.0, iso_z], F)
    wet = np.full((nz, 1, nx), wet_cm, F)
    dvp, d, sg, ii = _model(r80, A, B, C, idd, sig_ms)
    spots = np.array([[spot[0], spot[1], mu, 1]], F)
    D = (F(VSADX) + F(VSADY)) / F(2.0)
    g = O.geom(wet.shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, D)
    dose, _ = O.pencil(g, VSADX, VSADY, wet, spots, dvp, d, sg, ii)
    x = (np.arange(nx) * s)[None, :]
    z = (np.arange(nz) * s - iso_z)[:, None]
    dist = np.sqrt(x * x + float(D) ** 2 + z * z)   # |S - p| with S = (0, -D, 0)
    return dose[:, 0, :], x, z, dist


def _row_dose(wet_cm, r80, A, B, C, idd, sig_ms, mu=1.5, n=61, s=1.0):
    dose, x, _, dist = _plane_dose(wet_cm, r80, A, B, C, idd, sig_ms, mu, nx=n, s=s)
    return dose[0], x[0], dist[0]


def pin_sigma_air_argument():
    # R80 = 5 mm and w = 3 mm keep the halo weight at 0: the dose is a single
    # Gaussian of width sigma(p) = A d^2 + B d + C + sigma_ms
    r80, w = 5.0, 0.3
    for A, B, C, sms in ((1e-6, 1e-3, 2.0, 0.0), (-4e-7, 2.5e-3, 1.0, 1.5), (2e-6, -1e-3, 4.0, 0.7)):
        A, B, C, sms = (float(F(v)) for v in (A, B, C, sms))   # the tables hold float32
        idd, mu = 2.0, 1.5
        dose, x, z, dist = _plane_dose(w, r80, A, B, C, idd, sms, mu, nx=41, nz=21, s=1.5,
                                       iso_z=15.0)
        da = dist - float(F(F(w) * F(10.0))) + 0.7 * float(F(r80))
        sig = A * da * da + B * da + C + sms
        # on the axis (x = z = 0): dose = mu idd / (2 pi sigma^2)
        sig0 = math.sqrt(mu * idd / (2 * math.pi * dose[10, 0]))
        assert abs(sig0 - sig[10, 0]) <= 1e-7 * sig[10, 0], (A, B, C, sms, sig0, sig[10, 0])
        # off the axis: the exponent uses each voxel's own sigma
        d2 = x * x + z * z
        expect = mu * idd / (2 * math.pi * sig ** 2) * np.exp(-d2 / (2 * sig ** 2))
        m = expect > 1e-12 * expect.max()
        np.testing.assert_allclose(dose[m], expect[m], rtol=2e-6)


def pin_cax_distance_divergent():
    # a spot far off the central axis, voxels spread along the beam: the
    # squared distance is to the divergent spot axis through (x0, y0, 0) with
    # direction (x0 / VSADX, y0 / VSADY, -1), from a cross product
    C, idd, mu = 20.0, 2.0, 1.0
    x0, y0 = 40.0, 10.0
    shape = (1, 41, 41)
    s = 2.5
    iso = np.array([50.0, 50.0, -10.0], F)       # the plane sits at image z = 10
    wet = np.full(shape, 0.3, F)
    dvp, d, sg, ii = _model(5.0, 0.0, 0.0, C, idd, 0.0)
    D = (F(VSADX) + F(VSADY)) / F(2.0)
    g = O.geom(shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, D)
    dose, _ = O.pencil(g, VSADX, VSADY, wet, np.array([[x0, y0, mu, 1]], F), dvp, d, sg, ii)
    j, i = np.meshgrid(np.arange(41), np.arange(41), indexing="ij")
    px, py, pz = i * s - 50.0, j * s - 50.0, np.full(i.shape, 10.0)
    h = np.stack([px, pz, -py], -1)
    tau = np.array([x0 / VSADX, y0 / VSADY, -1.0])
    r = h - np.array([x0, y0, 0.0])
    d2 = (np.cross(r, tau) ** 2).sum(-1) / (tau @ tau)
    expect = mu * idd / (2 * np.pi * C * C) * np.exp(-d2 / (2 * C * C))
    np.testing.assert_allclose(dose[0], expect, rtol=2e-6)


def pin_halo_composition():
    # A = B = 0, sigma_ms = 0: the primary width is C everywhere on the row;
    # the tail (d >= 30 mm, primary below e^-50) is the halo Gaussian alone
    C, idd, mu = 3.0, 2.0, 1.0
    for r80, w in ((150.0, 10.0), (120.0, 4.0), (90.0, 8.5)):
        dose, x, _ = _row_dose(w, r80, 0.0, 0.0, C, idd, 0.0, mu, n=81)
        i1, i2 = 30, 60
        ratio = dose[i1] / dose[i2]
        sig_ht2 = (x[i2] ** 2 - x[i1] ** 2) / (2.0 * math.log(ratio))
        om = dose[i2] * 2 * math.pi * sig_ht2 / (mu * idd) * math.exp(x[i2] ** 2 / (2 * sig_ht2))
        # halo sigma and weight of nuclearHalo at this depth (empirical forms)
        wc = min(max(w * 10.0, 0.1), r80 - 0.1)
        sh = (2.85 + 0.0014 * r80 * math.log(wc + 3.0) + 0.06 * wc - 7.4e-5 * wc * wc
              - 0.22 * r80 / (wc - r80 - 5.0) ** 2)
        hw = (0.052 * math.log(1.13 + wc / (11.2 - 0.023 * r80))
              + 0.35 * (0.0017 * r80 * r80 - r80) / ((r80 + 3.0) ** 2 - wc * wc)
              - 1.61e-9 * wc * (r80 + 3.0) ** 2)
        sh, hw = max(sh, 0.0), min(max(hw, 0.0), 0.9)
        assert hw > 0.01
        assert abs(sig_ht2 - (C * C + sh * sh)) <= 1e-5 * sig_ht2, (r80, w, sig_ht2, C * C + sh * sh)
        assert abs(om - hw) <= 1e-5 * hw, (r80, w, om, hw)
        # axis: (1 - w) P + w H with both Gaussians normalised
        ax = mu * idd * ((1 - hw) / (2 * math.pi * C * C) + hw / (2 * math.pi * sig_ht2))
        assert abs(dose[0] - ax) <= 1e-6 * ax, (dose[0], ax)


PINS = [pin_rotation_round_trip_and_norm, pin_source_on_head_axis,
        pin_rotation_equals_library_matrices, pin_collimator_turns_about_central_axis,
        pin_tilt_maps_axis_onto_voxel_direction, pin_tilt_identity_on_central_axis,
        pin_sigma_air_argument, pin_cax_distance_divergent, pin_halo_composition]


@pytest.mark.parametrize("pin", PINS, ids=lambda f: f.__name__)
def test_pin(pin):
    pin()


def test_pin_cone_directions(photon_model):
    pin_cone_directions(photon_model.kernel)
