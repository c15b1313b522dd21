"""GPU tests of the plan layer (dosecuda_amd.plan).

1. The reference's two phantom scenarios (tests/test_phantom_impt.py and
   test_phantom_imrt.py: the default 138^3 cube, a 98-spot ring of energies,
   two square photon fields) through computeIMPTPlan / computeIMRTPlan equal,
   bit for bit, the reference's own loop over the drop-in ``dose_kernels``
   calls with its numpy accumulation (plan_impt.py:148-160,
   plan_imrt.py:140-158 with numpy < 2's float32 arithmetic).
2. The WET split (computeIMPTWET + computeIMPTPlanFromWET) gives the same
   per-field doses, and the NRRD exports read back as the computed arrays.
3. On a small cube the plan sums agree with the FP64 oracle's
   Σ outputFactor·dose within 1e-4·Dmax.
"""
import numpy as np
import pytest

from dosecuda_amd import dose_kernels, nrrd, plan as P
from oracle import oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.fixture(autouse=True)
def _need_gpu(gpu_ok):
    return gpu_ok


def _ring_beam(label="0", n_spots=98, radius=100.0, mu=0.2, e0=0, gantry=0.0, couch=0.0):
    """tests/test_phantom_impt.py:21-28: a circle of spots, energy id rising
    with the angle."""
    b = P.IMPTBeam()
    b.dicom_rangeshifter_label = label
    b.gantry_angle, b.couch_angle = gantry, couch
    for k in range(n_spots):
        theta = 2.0 * 3.14159 * k / n_spots
        b.addSingleSpot(radius * np.cos(theta), radius * np.sin(theta), mu, e0 + k)
    return b


def _impt_reference_loop(grid, plan):
    """plan_impt.py:124-160 as written, over the drop-in entry points."""
    rlsp = P.VolumeObject()
    rlsp.voxel_data = np.array(grid.RLSPFromHU(plan.machine_name), dtype=np.single)
    rlsp.origin = np.array(grid.origin, dtype=np.single)
    rlsp.spacing = np.array(grid.spacing, dtype=np.single)
    dose = np.zeros(grid.size, dtype=np.single)
    beams = []
    for beam in plan.beam_list:
        i = list(plan.dicom_rangeshifter_label.astype(str)).index(beam.dicom_rangeshifter_label)
        model = plan.beam_models[i]
        wet = dose_kernels.proton_raytrace_cuda(model, rlsp, beam, 0)
        w = P.VolumeObject()
        w.voxel_data = np.array(wet, dtype=np.single)
        w.origin, w.spacing = rlsp.origin, rlsp.spacing
        bd = dose_kernels.proton_spot_cuda(model, rlsp, w, beam, 0)
        beams.append(bd * plan.n_fractions)
        dose += bd
    dose *= plan.n_fractions
    return dose, beams


def _imrt_reference_loop(grid, plan):
    """plan_imrt.py:126-158; ``cp_dose * output_factor`` in float32 as
    numpy < 2 evaluates a float32 array times a float64 scalar."""
    vol = P.VolumeObject()
    vol.voxel_data = np.array(grid.DensityFromHU(plan.machine_name), dtype=np.single)
    vol.origin = np.array(grid.origin, dtype=np.single)
    vol.spacing = np.array(grid.spacing, dtype=np.single)
    dose = np.zeros(grid.size, dtype=np.single)
    beams = []
    for beam in plan.beam_list:
        bd = np.zeros(vol.voxel_data.shape, dtype=np.single)
        i = list(plan.dicom_energy_label.astype(str)).index(beam.dicom_energy_label)
        model = plan.beam_models[i]
        for cp in beam.cp_list:
            of = model.outputFactor(cp)
            cp_dose = dose_kernels.photon_dose_cuda(model, vol, cp, 0)
            bd += cp_dose * np.float32(of)
        beams.append(bd)
        dose += bd
    dose *= plan.n_fractions
    return dose, beams


@pytest.mark.parametrize("n_fractions", [1, 3])
def test_impt_phantom_plan_equals_reference_loop(n_fractions):
    print(f"[case] impt cube plan n_fractions={n_fractions}", flush=True)
    g = P.IMPTDoseGrid()
    g.createCubePhantom()
    p = P.IMPTPlan()
    p.n_fractions = n_fractions
    p.addBeam(_ring_beam())
    if n_fractions > 1:
        # a second field through the range shifter, rotated gantry and couch
        p.addBeam(_ring_beam("4", n_spots=40, radius=60.0, mu=0.5, e0=30, gantry=90.0,
                             couch=10.0))
    g.computeIMPTPlan(p)
    ref, ref_beams = _impt_reference_loop(g, p)
    assert g.dose.dtype == F and g.dose.shape == (138, 138, 138)
    assert float(g.dose.max()) > 0.0
    assert np.array_equal(g.dose, ref)
    assert len(g.beam_doses) == len(ref_beams)
    for a, b in zip(g.beam_doses, ref_beams):
        assert np.array_equal(a, b)
    assert np.array_equal(g.dose_device.cpu().numpy(), g.dose)


@pytest.mark.parametrize("n_fractions", [1, 2])
def test_imrt_phantom_plan_equals_reference_loop(n_fractions):
    print(f"[case] imrt cube plan n_fractions={n_fractions}", flush=True)
    g = P.IMRTDoseGrid()
    g.createCubePhantom()
    p = P.IMRTPlan("VarianTrueBeamHF")
    p.n_fractions = n_fractions
    p.addSquareField()
    p.addSquareField('6', 15.0, 15.0, mu=50, gantry_angle=90.0, collimator_angle=20.0)
    if n_fractions > 1:
        # a multi-control-point beam: three apertures at three gantry angles
        b = P.IMRTBeam()
        b.dicom_energy_label = '10'
        for k, (dx, dy, ga) in enumerate(((4, 6, 200.0), (8, 5, 210.0), (12, 12, 220.0))):
            q = P.IMRTPlan()
            q.addSquareField('10', dx, dy, mu=20 + 10 * k, gantry_angle=ga, table_angle=5.0)
            b.addControlPoint(q.beam_list[0].cp_list[0])
        p.addBeam(b)
    g.computeIMRTPlan(p)
    ref, ref_beams = _imrt_reference_loop(g, p)
    assert float(g.dose.max()) > 0.0
    assert np.array_equal(g.dose, ref)
    for a, b in zip(g.beam_doses, ref_beams):
        assert np.array_equal(a, b)


def test_impt_wet_split_and_exports(tmp_path):
    """plan_impt.py:162-242: the WET split gives each field's dose, and the
    NRRD exports hold the computed arrays."""
    print("[case] impt wet split + exports", flush=True)
    g = P.IMPTDoseGrid()
    g.createCubePhantom(size=[60, 64, 70], spacing=3.0)
    p = P.IMPTPlan()
    p.n_fractions = 2
    p.addBeam(_ring_beam(n_spots=24, radius=40.0, mu=1.0, e0=20))
    p.addBeam(_ring_beam("4", n_spots=12, radius=30.0, mu=1.0, e0=40, gantry=270.0))
    g.computeIMPTPlan(p)
    wets = g.computeIMPTWET(p)
    assert len(wets) == 2
    for i, (w, beam) in enumerate(zip(wets, p.beam_list)):
        assert w.voxel_data.dtype == F and w.voxel_data.shape == (60, 64, 70)
        d = g.computeIMPTPlanFromWET(p, w, beam)
        assert np.array_equal(d * p.n_fractions, g.beam_doses[i])
    g.writeWETNRRD(p, str(tmp_path / "wet.nrrd"))
    for i, w in enumerate(wets):
        a, o, s = nrrd.read(str(tmp_path / ("wet_beam%02d.nrrd" % (i + 1))))
        assert np.array_equal(a, w.voxel_data)
        assert np.array_equal(o, g.origin) and np.array_equal(s, g.spacing)
    g.writeDoseNRRD(str(tmp_path / "dose.nrrd"), dose_type="EFFECTIVE")
    assert np.array_equal(nrrd.read(str(tmp_path / "dose.nrrd"))[0], (g.dose * 1.1).astype(F))
    g.writeDoseNRRD(str(tmp_path / "dose.nrrd"), individual_beams=True, dose_type="PHYSICAL")
    for i in range(2):
        a = nrrd.read(str(tmp_path / ("dose_beam%02d.nrrd" % (i + 1))))[0]
        assert np.array_equal(a, g.beam_doses[i])


def test_imrt_plan_matches_oracle_small_cube():
    """Σ_beams Σ_cp outputFactor·D_cp of the FP64 oracle, times the fraction
    count, against the device plan: within 1e-4·Dmax."""
    print("[case] imrt plan vs oracle", flush=True)
    g = P.IMRTDoseGrid()
    g.createCubePhantom(size=[26, 30, 34], spacing=3.0)
    p = P.IMRTPlan()
    p.n_fractions = 2
    p.addSquareField('6', 6.0, 6.0, mu=100)
    p.addSquareField('10', 5.0, 4.0, mu=60, gantry_angle=90.0, collimator_angle=20.0)
    g.computeIMRTPlan(p)
    rho = g.DensityFromHU(p.machine_name)
    ref = np.zeros(g.HU.shape)
    for beam in p.beam_list:
        m = p.beam_models[list(p.dicom_energy_label.astype(str)).index(beam.dicom_energy_label)]
        for cp in beam.cp_list:
            d = O.photon_dose(rho, 3.0, g.origin, cp.iso, float(cp.mu), cp.ga, cp.ca, cp.ta, m,
                              cp.mlc)
            ref += m.outputFactor(cp) * d
    ref *= p.n_fractions
    mx, rms = rel_err(g.dose, ref)
    print(f"[parity] imrt plan: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= 1e-4


def test_impt_plan_matches_oracle_small_cube():
    """Each field's smoothed WET against the oracle (1e-5 of its maximum),
    then the oracle's pencil-beam dose on the engine's float32 WET, summed
    and scaled as the plan does (1e-4·Dmax)."""
    print("[case] impt plan vs oracle", flush=True)
    g = P.IMPTDoseGrid()
    g.createCubePhantom(size=[30, 32, 36], spacing=3.0)
    p = P.IMPTPlan()
    p.n_fractions = 2
    p.addBeam(_ring_beam(n_spots=8, radius=12.0, mu=1.0, e0=10))
    p.addBeam(_ring_beam("4", n_spots=6, radius=9.0, mu=1.0, e0=25, gantry=90.0, couch=15.0))
    g.computeIMPTPlan(p)
    wets = g.computeIMPTWET(p)
    ref = np.zeros(g.HU.shape)
    for beam, w in zip(p.beam_list, wets):
        m = p.beam_models[list(p.dicom_rangeshifter_label.astype(str))
                          .index(beam.dicom_rangeshifter_label)]
        _, sm, _ = O.proton_wet(g.RLSP, 3.0, g.origin, beam.iso, beam.gantry_angle,
                                beam.couch_angle, m.VSADX, m.VSADY)
        assert rel_err(w.voxel_data, sm)[0] <= 1e-5
        d, _ = O.proton_dose(w.voxel_data, 3.0, g.origin, beam.iso, beam.gantry_angle,
                             beam.couch_angle, m.VSADX, m.VSADY, beam.spot_list,
                             m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
        ref += d
    ref *= p.n_fractions
    mx, rms = rel_err(g.dose, ref)
    print(f"[parity] impt plan: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert float(ref.max()) > 0.0
    assert mx <= 1e-4
