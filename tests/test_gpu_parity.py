"""GPU parity: every stage of the HIP engine against the FP64 oracle on the
same inputs, then the drop-in entry points end to end.

Tolerance: max |GPU - oracle| <= 1e-4 * Dmax of the oracle for
doses; the WET stages are held to 1e-5 of their maximum.  Each test prints
its max and RMS error relative to the oracle's maximum.
Stage tests feed both sides the same float32 input, so a float32 decision
(range cut, TERMA <= 0.01, smoothing radius) cannot differ between them.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu
F = np.float32
TOL_DOSE = 1e-4
TOL_WET = 1e-5


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _cuda(t, a):
    return t.from_numpy(np.ascontiguousarray(a, dtype=F)).cuda()


def _report(name, got, ref, tol):
    mx, rms = rel_err(got, ref)
    print(f"[parity] {name}: max {mx:.3e} rms {rms:.3e} (x Dmax={np.abs(ref).max():.4g}) tol {tol:g}",
          flush=True)
    assert np.all(np.isfinite(got)), name
    assert mx <= tol, name
    return mx, rms


def _hetero(shape, seed, lo=0.2, hi=1.8):
    rng = np.random.default_rng(seed)
    base = rng.uniform(lo, hi, (shape[0] // 4 + 2, shape[1] // 4 + 2, shape[2] // 4 + 2))
    # blocky heterogeneity: nearest-neighbour upsampling of a coarse field
    idx = [np.minimum(np.arange(n) // 4, base.shape[a] - 1) for a, n in enumerate(shape)]
    return base[np.ix_(*idx)].astype(F)


RT_CASES = [
    ((20, 24, 28), 2.5, 0.0, 0.0, 0.0),
    ((24, 20, 16), 2.0, 57.0, 10.0, 0.0012),
    ((18, 22, 26), 3.0, 250.0, 340.0, 0.0),
]


@pytest.mark.parametrize("shape, s, ga, ta, border", RT_CASES)
def test_raytrace_stage(shape, s, ga, ta, border):
    t = _torch()
    from dosecuda_amd import engine
    rho = _hetero(shape, 11)
    iso = np.array([(shape[2] - 1) * s / 2 + 1.3, (shape[1] - 1) * s / 2 - 0.7, (shape[0] - 1) * s / 2], F)
    origin = np.array([-3.0, 2.0, 1.0], F)
    print(f"[case] raytrace shape={shape} s={s} ga={ga} ta={ta}", flush=True)
    got = engine.stage_raytrace(_cuda(t, rho), s, origin, iso, ga, ta, 1635.0, border).cpu().numpy()
    g = O.geom(shape, s, origin, iso, ga, ta, 0.0, 1635.0)
    ref, _ = O.raytrace(g, rho, border)
    _report("raytrace", got, ref, TOL_WET)


@pytest.mark.parametrize("shape, s, ga, ta, border", RT_CASES)
def test_smooth_stage(shape, s, ga, ta, border):
    t = _torch()
    from dosecuda_amd import engine
    rng = np.random.default_rng(12)
    wet = (rng.uniform(0.0, 2.0, shape) + np.linspace(0, 3, shape[1])[None, :, None]).astype(F)
    iso = np.array([(shape[2] - 1) * s / 2, (shape[1] - 1) * s / 2, (shape[0] - 1) * s / 2], F)
    print(f"[case] smooth shape={shape} ga={ga} ta={ta}", flush=True)
    got = engine.stage_smooth(_cuda(t, wet), s, np.zeros(3), iso, ga, ta, 1340.0, 1930.0).cpu().numpy()
    g = O.geom(shape, s, np.zeros(3), iso, ga, ta, 0.0, (F(1340.0) + F(1930.0)) / F(2.0))
    ref = O.smooth(g, wet)
    _report("smooth", got, ref, TOL_WET)


def _impt_case(proton_model, shape=(36, 44, 40), s=2.5, n_layers=4, side=8, seed=21, ga=0.0):
    from dosecuda_amd import phantoms
    rng = np.random.default_rng(seed)
    rho = np.full(shape, 1.00653, F)
    d0 = int(rng.integers(8, 16))
    rho[:, d0:d0 + 6, :] = 1.5562
    iso = np.array([(shape[2] - 1) * s / 2, (shape[1] - 1) * s / 2, (shape[0] - 1) * s / 2], F)
    spots, _ = phantoms.impt_spot_grid(proton_model, seed=seed, n_side=side, pitch=5.0,
                                       n_layers=n_layers, e_lo=70.0, e_hi=110.0)
    return rho, iso, spots


@pytest.mark.parametrize("ga", [0.0, 135.0])
def test_pencil_stage(proton_model, ga):
    t = _torch()
    from dosecuda_amd import engine
    m = proton_model
    rho, iso, spots = _impt_case(m, ga=ga)
    shape, s = rho.shape, 2.5
    g = O.geom(shape, s, np.zeros(3), iso, ga, 0.0, 0.0, (F(m.VSADX) + F(m.VSADY)) / F(2.0))
    raw, _ = O.raytrace(g, rho, 0.0)
    wet = O.smooth(g, raw.astype(F)).astype(F)
    print(f"[case] pencil shape={shape} ga={ga} spots={len(spots)}", flush=True)
    plan = engine.ProtonPlan(0, shape, s, np.zeros(3), iso, ga, 0.0, m, spots)
    got = plan.dose(_cuda(t, wet)).cpu().numpy()
    ref, pairs = O.pencil(g, m.VSADX, m.VSADY, wet, spots, m.divergence_params, m.lut_depths,
                          m.lut_sigmas, m.lut_idds)
    assert pairs == plan.pairs(wet)
    _report("pencil", got, ref, TOL_DOSE)
    plan.close()


def _imrt_case(photon_model, shape=(28, 32, 30), s=2.5, seed=31):
    rng = np.random.default_rng(seed)
    rho = np.full(shape, 1.00653, F)
    c = np.array(shape) // 2 + rng.integers(-3, 4, 3)
    k, j, i = np.ogrid[:shape[0], :shape[1], :shape[2]]
    lung = ((k - c[0]) / 5.0) ** 2 + ((j - c[1]) / 6.0) ** 2 + ((i - c[2]) / 4.0) ** 2 <= 1
    rho[np.broadcast_to(lung, shape)] = 0.3037
    bone = (j - c[1] - 4) ** 2 + (i - c[2] + 5) ** 2 <= 9
    rho[np.broadcast_to(bone, shape)] = 1.5562
    iso = np.array([(shape[2] - 1) * s / 2, (shape[1] - 1) * s / 2, (shape[0] - 1) * s / 2], F)
    return rho, iso


@pytest.mark.parametrize("ga, ca, ta", [(0.0, 0.0, 0.0), (51.43, 20.0, 0.0), (257.0, 0.0, 10.0)])
def test_terma_stage(photon_model, ga, ca, ta):
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    rho, iso = _imrt_case(pm)
    shape, s = rho.shape, 2.5
    mlc = square_field_mlc(pm, 5.0, 4.0)
    g = O.geom(shape, s, np.zeros(3), iso, ga, ta, ca, F(pm.primary_source_distance))
    raw, _ = O.raytrace(g, rho, float(F(0.0012)))
    wet = raw.astype(F)
    print(f"[case] terma shape={shape} ga={ga} ca={ca} ta={ta}", flush=True)
    tg, eg = engine.stage_terma(_cuda(t, wet), s, np.zeros(3), iso, 100.0, ga, ca, ta, pm, mlc)
    tr, er = O.terma(g, pm, mlc, wet)
    _report("terma", tg.cpu().numpy(), tr, 1e-5)
    _report("electron", eg.cpu().numpy(), er, 1e-5)


@pytest.mark.parametrize("ga, ca", [(0.0, 0.0), (77.0, 25.0)])
def test_terma_pair_window_is_exact(photon_model, ga, ca):
    """The y-sorted leaf pairs are summed over a binary-searched window; a
    far-away open pair placed first breaks the sort and forces the full scan.
    It contributes exactly 0 everywhere, so both sums must agree bit for bit."""
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    rho, iso = _imrt_case(pm)
    s = 2.5
    rng = np.random.default_rng(3)
    mlc = square_field_mlc(pm, 8.0, 6.0).copy()
    mlc[:, 0] += rng.uniform(-8, 8, mlc.shape[0]).astype(F)        # irregular leaf edges
    mlc[:, 1] = np.maximum(mlc[:, 1] + rng.uniform(-8, 8, mlc.shape[0]).astype(F), mlc[:, 0])
    far = np.array([[-5.0, 5.0, 1000.0, 5.0]], F)                   # open, 1 m off axis
    wet = _hetero(rho.shape, 4, 0.0, 2.0)
    print(f"[case] terma pair window ga={ga} ca={ca}", flush=True)
    a, ea = engine.stage_terma(_cuda(t, wet), s, np.zeros(3), iso, 100.0, ga, ca, 0.0, pm, mlc)
    b, eb = engine.stage_terma(_cuda(t, wet), s, np.zeros(3), iso, 100.0, ga, ca, 0.0, pm,
                               np.vstack([far, mlc]))
    assert float(a.max()) > 0.0
    assert t.equal(a, b) and t.equal(ea, eb)


def _modulated_mlc(pm, seed):
    """Irregular aperture: random leaf tips, closed pairs at random positions
    inside the field (equal tips away from the axis), and a run of pairs with
    equal tips, so TERMA's shared y-edge erf, shared x factor and closed-pair
    shortcuts all take both branches."""
    from dosecuda_amd.beam_models import square_field_mlc
    rng = np.random.default_rng(seed)
    mlc = square_field_mlc(pm, 12.0, 12.0).copy()
    n = mlc.shape[0]
    x1 = rng.uniform(-50.0, 10.0, n).astype(F)
    mlc[:, 0] = x1
    mlc[:, 1] = x1 + rng.uniform(0.0, 60.0, n).astype(F)
    closed = rng.choice(n, n // 5, replace=False)
    mlc[closed, 0] = rng.uniform(-40.0, 40.0, closed.size).astype(F)
    mlc[closed, 1] = mlc[closed, 0]
    mlc[n // 2 - 5:n // 2 + 5, 0] = -30.0
    mlc[n // 2 - 5:n // 2 + 5, 1] = 25.0
    return mlc


@pytest.mark.parametrize("ga, ca, shuffle", [(0.0, 0.0, False), (51.43, 20.0, False), (200.0, 0.0, True)])
def test_terma_modulated_mlc(photon_model, ga, ca, shuffle):
    """TERMA of an irregular MLC aperture against the FP64 oracle; shuffled
    pair order takes the full scan instead of the y-sorted window."""
    t = _torch()
    from dosecuda_amd import engine
    pm = photon_model
    rho, iso = _imrt_case(pm)
    shape, s = rho.shape, 2.5
    mlc = _modulated_mlc(pm, 41)
    if shuffle:
        mlc = mlc[np.random.default_rng(42).permutation(mlc.shape[0])]
    g = O.geom(shape, s, np.zeros(3), iso, ga, 0.0, ca, F(pm.primary_source_distance))
    raw, _ = O.raytrace(g, rho, float(F(0.0012)))
    wet = raw.astype(F)
    print(f"[case] terma modulated mlc ga={ga} ca={ca} shuffle={shuffle}", flush=True)
    tg, eg = engine.stage_terma(_cuda(t, wet), s, np.zeros(3), iso, 100.0, ga, ca, 0.0, pm, mlc)
    tr, er = O.terma(g, pm, mlc, wet)
    assert float(tr.max()) > 0.0
    _report("terma modulated", tg.cpu().numpy(), tr, 1e-5)
    _report("electron modulated", eg.cpu().numpy(), er, 1e-5)


@pytest.mark.parametrize("ga, ca, ta", [(0.0, 0.0, 0.0), (102.86, 30.0, 0.0), (308.57, 0.0, 20.0)])
def test_ccc_stage(photon_model, ga, ca, ta):
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    rho, iso = _imrt_case(pm)
    shape, s = rho.shape, 2.5
    mlc = square_field_mlc(pm, 5.0, 4.0)
    g = O.geom(shape, s, np.zeros(3), iso, ga, ta, ca, F(pm.primary_source_distance))
    raw, _ = O.raytrace(g, rho, float(F(0.0012)))
    tr, er = O.terma(g, pm, mlc, raw.astype(F))
    t32, e32 = tr.astype(F), er.astype(F)
    print(f"[case] ccc shape={shape} ga={ga} ca={ca} ta={ta}", flush=True)
    got = engine.stage_ccc(_cuda(t, t32), _cuda(t, rho), _cuda(t, e32), s, np.zeros(3), iso,
                           100.0, ga, ca, ta, pm).cpu().numpy()
    ref, n = O.ccc(g, pm, 100.0, t32, rho, e32)
    assert n > 0
    _report("ccc", got, ref, TOL_DOSE)


def test_dropin_impt_end_to_end(proton_model):
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import ProtonField, VolumeObject
    m = proton_model
    rho, iso, spots = _impt_case(m, shape=(30, 40, 34), n_layers=3, side=6)
    s = 2.5
    vol = VolumeObject(rho, origin=(0, 0, 0), spacing=(s, s, s))
    beam = ProtonField(iso, 0.0, 0.0, spots)
    print("[case] drop-in impt", flush=True)
    wet = dose_kernels.proton_raytrace_cuda(m, vol, beam, 0)
    assert wet.dtype == np.float32 and wet.shape == rho.shape
    raw, sm, _ = O.proton_wet(rho, s, np.zeros(3), iso, 0.0, 0.0, m.VSADX, m.VSADY)
    _report("dropin wet", wet, sm, TOL_WET)
    wv = VolumeObject(wet, origin=(0, 0, 0), spacing=(s, s, s))
    dose = dose_kernels.proton_spot_cuda(m, vol, wv, beam, 0)
    ref, _ = O.proton_dose(wet, s, np.zeros(3), iso, 0.0, 0.0, m.VSADX, m.VSADY, spots,
                           m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
    _report("dropin impt dose", dose, ref, TOL_DOSE)


def test_dropin_photon_end_to_end(photon_model):
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import PhotonControlPoint, VolumeObject, square_field_mlc
    pm = photon_model
    rho, iso = _imrt_case(pm, shape=(26, 30, 28))
    s = 2.5
    vol = VolumeObject(rho, origin=(0, 0, 0), spacing=(s, s, s))
    cp = PhotonControlPoint(iso, 100.0, square_field_mlc(pm, 6.0, 5.0), 30.0, 15.0, 0.0)
    print("[case] drop-in photon", flush=True)
    dose = dose_kernels.photon_dose_cuda(pm, vol, cp, 0)
    assert dose.dtype == np.float32 and dose.shape == rho.shape
    ref = O.photon_dose(rho, s, np.zeros(3), iso, 100.0, 30.0, 15.0, 0.0, pm, cp.mlc)
    _report("dropin photon dose", dose, ref, TOL_DOSE)


def test_dropin_photon_modulated_mlc(photon_model):
    """An irregular aperture (closed pairs inside the field, equal-tip runs)
    through the drop-in, with collimator and couch rotation, against the
    FP64 oracle's whole beam."""
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import PhotonControlPoint, VolumeObject
    pm = photon_model
    rho, iso = _imrt_case(pm, shape=(26, 30, 28), seed=43)
    s = 2.5
    vol = VolumeObject(rho, origin=(0, 0, 0), spacing=(s, s, s))
    cp = PhotonControlPoint(iso, 100.0, _modulated_mlc(pm, 44), 320.0, 35.0, 10.0)
    print("[case] drop-in photon modulated mlc", flush=True)
    dose = dose_kernels.photon_dose_cuda(pm, vol, cp, 0)
    ref = O.photon_dose(rho, s, np.zeros(3), iso, 100.0, 320.0, 35.0, 10.0, pm, cp.mlc)
    assert float(ref.max()) > 0.0
    _report("dropin photon modulated", dose, ref, TOL_DOSE)


def test_invalid_gpu_id_raises(proton_model):
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import ProtonField, VolumeObject
    vol = VolumeObject(np.ones((4, 4, 4), F), spacing=(2, 2, 2))
    beam = ProtonField(np.zeros(3), 0.0, 0.0, [[0, 0, 1, 10]])
    with pytest.raises(RuntimeError, match="CUDA error: "):
        dose_kernels.proton_raytrace_cuda(proton_model, vol, beam, 97)


def test_full_impt_field_100cubed(proton_model):
    """Plain IMPT at full size: 100^3 @ 2.5 mm, water with a seeded 2 cm bone
    slab, 20x20 spots x 10 layers (100-160 MeV), unit MU, gantry 0."""
    t = _torch()
    from dosecuda_amd import engine, phantoms
    m = proton_model
    case = phantoms.impt_slab_case(seed=1001, n=100, spacing=2.5)
    rlsp = phantoms.to_rlsp(case)
    spots, _ = phantoms.impt_spot_grid(m, seed=1001)
    print(f"[case] impt 100^3 spots={len(spots)}", flush=True)
    d = _cuda(t, rlsp)
    raw = t.empty_like(d)
    wet = engine.proton_wet(d, case.spacing, case.origin, case.iso, 0.0, 0.0, m.VSADX, m.VSADY,
                            raw_out=raw)
    plan = engine.ProtonPlan(0, rlsp.shape, case.spacing, case.origin, case.iso, 0.0, 0.0, m, spots)
    dose = plan.dose(wet).cpu().numpy()
    wet = wet.cpu().numpy()
    raw = raw.cpu().numpy()
    ref_raw, ref_sm, _ = O.proton_wet(rlsp, case.spacing, case.origin, case.iso, 0.0, 0.0,
                                      m.VSADX, m.VSADY)
    _report("impt100 raw wet", raw, ref_raw, TOL_WET)
    # smoothing decisions come from the GPU's own float32 raw WET
    g = O.geom(rlsp.shape, case.spacing, case.origin, case.iso, 0.0, 0.0, 0.0,
               (F(m.VSADX) + F(m.VSADY)) / F(2.0))
    _report("impt100 smoothed wet", wet, O.smooth(g, raw), TOL_WET)
    ref, pairs = O.pencil(g, m.VSADX, m.VSADY, wet, spots, m.divergence_params, m.lut_depths,
                          m.lut_sigmas, m.lut_idds)
    print(f"[parity] impt100 pairs {pairs}", flush=True)
    _report("impt100 dose", dose, ref, TOL_DOSE)
    plan.close()


def test_full_cccs_beam_128cubed(photon_model):
    """One oblique beam (gantry 51.43) of the plain CCCS plan on the 128^3
    lung/bone phantom, 10x10 cm field, each stage against the oracle."""
    t = _torch()
    from dosecuda_amd import engine, phantoms
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    case = phantoms.imrt_lung_bone_case(seed=1002, n=128, spacing=2.5)
    dens = phantoms.to_density(case)
    mlc = square_field_mlc(pm, 10.0, 10.0)
    ga = 360.0 / 7
    print("[case] cccs 128^3 gantry 51.43", flush=True)
    d = _cuda(t, dens)
    terma, el, wet = t.empty_like(d), t.empty_like(d), t.empty_like(d)
    dose = engine.photon_dose(d, case.spacing, case.origin, case.iso, 100.0, ga, 0.0, 0.0, pm, mlc,
                              terma_out=terma, electron_out=el, wet_out=wet)
    dose, terma, el, wet = (x.cpu().numpy() for x in (dose, terma, el, wet))
    g = O.geom(dens.shape, case.spacing, case.origin, case.iso, ga, 0.0, 0.0,
               F(pm.primary_source_distance))
    ref_wet, _ = O.raytrace(g, dens, float(F(0.0012)))
    _report("cccs128 wet", wet, ref_wet, TOL_WET)
    tr, er = O.terma(g, pm, mlc, wet)
    _report("cccs128 terma", terma, tr, 1e-5)
    _report("cccs128 electron", el, er, 1e-5)
    ref, n = O.ccc(g, pm, 100.0, terma, dens, el)
    print(f"[parity] cccs128 samples {n}", flush=True)
    _report("cccs128 dose", dose, ref, TOL_DOSE)


def test_full_cccs_plan_128cubed_every_beam(photon_model):
    """All 7 beams of the plain CCCS plan on the 128^3 lung/bone phantom
    (10x10 cm, gantry 360*b/7), each end to end against the oracle's own
    stage chain from the same density; the plan sum too.

    End to end, the engine's FP32 TERMA and the oracle's FP64 TERMA (rounded
    to FP32) can fall on opposite sides of the 0.01 early-out of cccKernel
    (IMRTClasses.cu:262-266) for a voxel whose TERMA is 0.01 to within
    rounding; such a voxel is all or nothing (0 against its full dose).
    Those decision flips are counted, each is checked to sit within 1e-5
    (relative) of the threshold, and they are left out of the error, which
    is otherwise held to 1e-4 of Dmax."""
    t = _torch()
    from dosecuda_amd import engine, phantoms
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    case = phantoms.imrt_lung_bone_case(seed=1002, n=128, spacing=2.5)
    dens = phantoms.to_density(case)
    mlc = square_field_mlc(pm, 10.0, 10.0)
    d = _cuda(t, dens)
    terma = t.empty_like(d)
    psum, rsum, keep = 0.0, 0.0, np.ones(dens.shape, bool)
    for ga in phantoms.imrt_beams(7):
        print(f"[case] cccs128 plan beam gantry={ga:.2f}", flush=True)
        got = engine.photon_dose(d, case.spacing, case.origin, case.iso, 100.0, ga, 0.0, 0.0, pm,
                                 mlc, terma_out=terma).cpu().numpy()
        tg = terma.cpu().numpy()
        ref, st = O.photon_dose(dens, case.spacing, case.origin, case.iso, 100.0, ga, 0.0, 0.0, pm, mlc,
                                return_stages=True)
        flip = (tg <= F(0.01)) != (st["terma32"] <= F(0.01))
        nf = int(flip.sum())
        print(f"[parity] cccs128 beam {ga:.2f}: {nf} TERMA early-out flips", flush=True)
        if nf:
            assert nf <= 16
            assert np.abs(st["terma"][flip] - 0.01).max() <= 1e-5 * 0.01
        diff = np.where(flip, 0.0, np.abs(got - ref))
        w = np.unravel_index(int(np.argmax(diff)), diff.shape)
        print(f"[parity] worst voxel {w}: engine {got[w]:.6g} oracle {ref[w]:.6g} TERMA engine {tg[w]:.8g} "
              f"oracle {st['terma'][w]:.8g}", flush=True)
        _report(f"cccs128 beam {ga:.2f}", np.where(flip, 0.0, got), np.where(flip, 0.0, ref), TOL_DOSE)
        keep &= ~flip
        psum = psum + got.astype(np.float64)
        rsum = rsum + ref
    _report("cccs128 plan sum", np.where(keep, psum, 0.0), np.where(keep, rsum, 0.0), TOL_DOSE)


@pytest.mark.parametrize("ga, ca, ta", [(0.0, 0.0, 0.0), (290.0, 15.0, 10.0)])
def test_dropin_photon_elekta_modulated(ga, ca, ta):
    """The second photon head: ElektaVersaSibAxesse 6 MV (80 x 5 mm leaf
    pairs, MLC plane at 643.2 mm, y jaws only) with an irregular aperture,
    through the drop-in, against the oracle's whole beam."""
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import PhotonBeamModel, PhotonControlPoint, VolumeObject
    pm = PhotonBeamModel("ElektaVersaSibAxesse", "6MV", "6")
    assert pm.n_mlc_pairs == 80 and pm.mlc_distance == 643.2
    rho, iso = _imrt_case(pm, shape=(26, 30, 28), seed=47)
    s = 2.5
    vol = VolumeObject(rho, origin=(0, 0, 0), spacing=(s, s, s))
    cp = PhotonControlPoint(iso, 100.0, _modulated_mlc(pm, 48), ga, ca, ta)
    print(f"[case] drop-in photon elekta ga={ga} ca={ca} ta={ta}", flush=True)
    dose = dose_kernels.photon_dose_cuda(pm, vol, cp, 0)
    ref = O.photon_dose(rho, s, np.zeros(3), iso, 100.0, ga, ca, ta, pm, cp.mlc)
    assert float(ref.max()) > 0.0
    _report("dropin photon elekta", dose, ref, TOL_DOSE)


def test_pencil_rs4_low_energy_layers():
    """Range shifter RS4's lowest layers: id 0 has R80 = -0.65 mm (every
    voxel is past the 1.1 * R80 cut, and the halo clamp's upper bound
    R80 - 0.1 lies below its lower bound), ids 1-3 have R80 of 1.5-3.5 mm,
    so only voxels at the surface receive dose.  Engine and oracle from the
    same WET; id 0 alone gives zero everywhere."""
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import ProtonBeamModel
    m = ProtonBeamModel(rangeshifter_folder="RS4", dicom_rangeshifter_label="4")
    assert m.divergence_params[0, 1] < 0.0
    shape, s = (20, 24, 22), 2.0
    rho = np.zeros(shape, F)
    rho[:, 6:, :] = 1.00653                       # air gap, then water from the upstream face
    iso = np.array([(shape[2] - 1) * s / 2, (shape[1] - 1) * s / 2, (shape[0] - 1) * s / 2], F)
    g = O.geom(shape, s, np.zeros(3), iso, 0.0, 0.0, 0.0, (F(m.VSADX) + F(m.VSADY)) / F(2.0))
    raw, _ = O.raytrace(g, rho, 0.0)
    wet = O.smooth(g, raw.astype(F)).astype(F)
    c = (np.arange(5) - 2) * 4.0
    xx, yy = np.meshgrid(c, c)
    spots = np.vstack([np.column_stack([xx.ravel(), yy.ravel(), np.ones(25), np.full(25, e)])
                       for e in (0, 1, 2, 3)]).astype(F)
    args = (m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
    for sel, name in ((spots, "rs4 ids 0-3"), (spots[spots[:, 3] == 0], "rs4 id 0")):
        print(f"[case] pencil {name}", flush=True)
        plan = engine.ProtonPlan(0, shape, s, np.zeros(3), iso, 0.0, 0.0, m, sel)
        got = plan.dose(_cuda(t, wet)).cpu().numpy()
        plan.close()
        ref, _ = O.pencil(g, m.VSADX, m.VSADY, wet, sel, *args)
        if name == "rs4 id 0":
            assert not np.any(ref) and not np.any(got)
        else:
            assert float(ref.max()) > 0.0
            _report(name, got, ref, TOL_DOSE)
