"""GPU: the tuners time every variant, find them all equal (bit for bit for
voxel tiles and LUT staging), and the tuned launches give the default
results."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_tune_proton_wet_and_pencil(proton_model):
    t = _torch()
    from dosecuda_amd import autotune, engine
    m = proton_model
    shape, s = (40, 52, 45), 2.5
    rlsp = t.ones(shape, device="cuda")
    rlsp[:, 20:26, :] = 1.7
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    autotune.clear()
    ref = engine.proton_wet(rlsp, s, np.zeros(3), iso, 30.0, 0.0, m.VSADX, m.VSADY)
    print("[case] tune proton wet", flush=True)
    r = autotune.tune_proton_wet(rlsp, s, np.zeros(3), iso, 30.0, 0.0, m, reps=3)
    print(f"[tune] {r}", flush=True)
    assert r["raytrace"] in autotune.TILES and min(r["raytrace_ms"]) > 0
    got = engine.proton_wet(rlsp, s, np.zeros(3), iso, 30.0, 0.0, m.VSADX, m.VSADY)
    assert t.equal(got, ref)
    rng = np.random.default_rng(2)
    spots = np.column_stack([rng.uniform(-20, 20, (80, 2)), np.ones(80), np.repeat([30, 36], 40)]).astype(np.float32)
    plan = engine.ProtonPlan(0, shape, s, np.zeros(3), iso, 30.0, 0.0, m, spots)
    d0 = plan.dose(ref)
    print("[case] tune pencil", flush=True)
    r = autotune.tune_pencil(plan, ref, reps=3)
    print(f"[tune] {r['tile']} lds={r['lut_in_lds']}", flush=True)
    assert t.equal(plan.dose(ref), d0)
    autotune.clear()


def test_tune_conv3d_variants_agree():
    t = _torch()
    from dosecuda_amd import autotune, surrogate as S
    autotune.clear()
    print("[case] tune conv3d", flush=True)
    r = autotune.tune_conv3d(1, 16, 16, 32, 32, 0, 0, 32, 3, reps=3)
    print(f"[tune] {r}", flush=True)
    assert r["best"] > 0 and all(c > 0 for c in r["codes"])      # every variant admitted
    g = t.Generator().manual_seed(1)
    x = t.randn((1, 16, 16, 32, 32), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((27, 32, 32), generator=g) * 0.1).cuda()
    wf = S.prep_weights(w, 0)
    _, tuned = S.conv3d(x, wf, None, 32, 3, False, want_bf16=False, want_f32=True)
    autotune.clear()
    _, dflt = S.conv3d(x, wf, None, 32, 3, False, want_bf16=False, want_f32=True)
    assert float((tuned - dflt).abs().max()) <= 1e-5 * float(dflt.abs().max())


def test_tune_influence(proton_model):
    t = _torch()
    from dosecuda_amd import autotune, engine
    m = proton_model
    shape, s = (30, 36, 32), 2.5
    rlsp = t.ones(shape, device="cuda")
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    wet = engine.proton_wet(rlsp, s, np.zeros(3), iso, 0.0, 0.0, m.VSADX, m.VSADY)
    rng = np.random.default_rng(4)
    spots = np.column_stack([rng.uniform(-15, 15, (50, 2)), np.ones(50), np.repeat([25, 33], 25)]).astype(np.float32)
    autotune.clear()
    print("[case] tune influence", flush=True)
    r = autotune.tune_influence(wet, shape, s, np.zeros(3), iso, 0.0, 0.0, m, spots, reps=2)
    print(f"[tune] {r['best']}", flush=True)
    f = engine.InfluenceMatrix(0, shape, s, np.zeros(3), iso, 0.0, 0.0, m, spots)
    assert f.count(wet) > 0
    autotune.clear()


@pytest.mark.parametrize("gantry", [0.0, 120.0])
def test_tune_influence_wet(proton_model, gantry):
    """The footprint WET's tile tuner stores its fastest tiles under the
    beam's geometry key and every tile gives the identical WET (the tuner
    raises otherwise); the key names the image axis of the source's largest
    component, as the library's own key does (y at gantry 0, x at 120)."""
    t = _torch()
    from dosecuda_amd import autotune, engine
    m = proton_model
    shape, s = (30, 36, 32), 2.5
    rlsp = t.ones(shape, device="cuda")
    rlsp[8:14, 10:20, 5:25] = 1.6
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    rng = np.random.default_rng(5)
    spots = np.column_stack([rng.uniform(-15, 15, (40, 2)), np.ones(40), np.repeat([25, 33], 20)]).astype(np.float32)
    autotune.clear()
    print(f"[case] tune influence wet gantry={gantry}", flush=True)
    r = autotune.tune_influence_wet(rlsp, shape, s, np.zeros(3), iso, gantry, 0.0, m, spots, reps=1)
    key = autotune.geom_key(shape, gantry, 0.0, (float(np.float32(m.VSADX)) + float(np.float32(m.VSADY))) / 2.0)
    for kernel in ("raytrace_box", "smooth_box"):
        assert autotune.TILES[autotune.get(kernel, key)] == r[kernel]["best"]
    assert key == "30,36,32," + ("1" if gantry == 0.0 else "0")
    autotune.clear()


def test_tune_photon(photon_model):
    t = _torch()
    from dosecuda_amd import autotune, engine
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    shape, s = (36, 44, 40), 2.5
    rho = t.full(shape, 1.00653, device="cuda")
    rho[10:20, 12:30, 8:22] = 0.3
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    mlc = square_field_mlc(pm, 7.0, 6.0)
    autotune.clear()
    ref = engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 20.0, 0.0, 0.0, pm, mlc)
    print("[case] tune photon", flush=True)
    r = autotune.tune_photon(rho, s, np.zeros(3), iso, 100.0, 20.0, 0.0, 0.0, pm, mlc, reps=3)
    print(f"[tune] {r}", flush=True)
    got = engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 20.0, 0.0, 0.0, pm, mlc)
    assert t.equal(got, ref)
    # the quad layout explicitly, also for the voxel-list path
    key = ",".join(map(str, shape)) + ",1"
    for layout in (0, 1):
        autotune.set_choice("ccc", key, layout)
        assert t.equal(engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 20.0, 0.0, 0.0, pm, mlc), ref)
    autotune.clear()
