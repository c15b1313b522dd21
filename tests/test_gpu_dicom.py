"""Plans that come in through DICOM compute the same doses as the plans they
encode (GPU): an RT Ion Plan written by IMPTPlan.writePlanDicom and read back
by readPlanDicom (plan_impt.py:307-643), and an RT Plan carrying the square
field of IMRTPlan.addSquareField read by IMRTPlan.readPlanDicom
(plan_imrt.py:273-351).  Bit-equal: the files carry the float32 spot data
and the MLC exactly, and the engine sorts spots by layer stably, so the
spot order within each layer is the same."""
import numpy as np
import pytest

from dosecuda_amd import dicom as D
from dosecuda_amd import plan as P
from tests.test_dicom import _bldp, _cp, _write_ct_series

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu(gpu_ok):
    return gpu_ok


def test_impt_plan_through_rt_ion_plan(tmp_path):
    print("[case] impt plan through RT Ion Plan", flush=True)
    ct = str(tmp_path / "ct")
    _write_ct_series(ct, np.zeros((3, 4, 5)), (0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    p = P.IMPTPlan()
    p.n_fractions = 3
    rng = np.random.default_rng(11)
    for label, gantry, couch in (("0", 0.0, 0.0), ("4", 90.0, 15.0)):
        b = P.IMPTBeam()
        b.dicom_rangeshifter_label = label
        b.gantry_angle, b.couch_angle = gantry, couch
        for e in rng.permutation(np.repeat([20, 35, 50], 6)):
            b.addSingleSpot(float(rng.uniform(-25, 25)), float(rng.uniform(-25, 25)),
                            float(rng.uniform(0.5, 1.5)), int(e))
        p.addBeam(b)
    path = str(tmp_path / "ion.dcm")
    p.writePlanDicom(ct, path)
    q = P.IMPTPlan()
    q.readPlanDicom(path)
    g1, g2 = P.IMPTDoseGrid(), P.IMPTDoseGrid()
    for g in (g1, g2):
        g.createCubePhantom(size=[40, 44, 48], spacing=3.0)
    g1.computeIMPTPlan(p)
    g2.computeIMPTPlan(q)
    assert float(g1.dose.max()) > 0.0
    assert np.array_equal(g1.dose, g2.dose)
    for a, b in zip(g1.beam_doses, g2.beam_doses):
        assert np.array_equal(a, b)


def test_imrt_square_field_through_rt_plan(tmp_path):
    print("[case] imrt square field through RT Plan", flush=True)
    p = P.IMRTPlan()
    p.n_fractions = 2
    p.addSquareField("6", 8.0, 6.0, mu=150, gantry_angle=30.0, collimator_angle=10.0, table_angle=5.0)
    cp = p.beam_list[0].cp_list[0]
    leaves = np.concatenate([cp.mlc[:, 0], cp.mlc[:, 1]])
    beam = D.Dataset()
    beam.BeamNumber = 1
    beam.TreatmentDeliveryType = "TREATMENT"
    beam.ControlPointSequence = [
        _cp(0, 0.0, NominalBeamEnergy="6", IsocenterPosition=[0.0, 0.0, 0.0], GantryAngle=30.0,
            BeamLimitingDeviceAngle=10.0, PatientSupportAngle=5.0,
            BeamLimitingDevicePositionSequence=_bldp([("ASYMX", (-40, 40)), ("ASYMY", (-30, 30)),
                                                      ("MLCX", leaves)])),
        _cp(1, 1.0),
    ]
    fg = D.Dataset()
    fg.NumberOfFractionsPlanned = 2
    rb = D.Dataset()
    rb.ReferencedBeamNumber = 1
    rb.BeamMeterset = 150
    fg.ReferencedBeamSequence = [rb]
    ds = D.Dataset()
    ds.SOPClassUID = D.RT_PLAN_STORAGE
    ds.SOPInstanceUID = D.generate_uid()
    ds.FractionGroupSequence = [fg]
    ds.BeamSequence = [beam]
    path = str(tmp_path / "plan.dcm")
    D.write(ds, path)
    q = P.IMRTPlan()
    q.readPlanDicom(path)
    assert np.array_equal(q.beam_list[0].cp_list[0].mlc, cp.mlc)
    g1, g2 = P.IMRTDoseGrid(), P.IMRTDoseGrid()
    for g in (g1, g2):
        g.createCubePhantom(size=[36, 40, 44], spacing=3.0)
    g1.computeIMRTPlan(p)
    g2.computeIMRTPlan(q)
    assert float(g1.dose.max()) > 0.0
    assert np.array_equal(g1.dose, g2.dose)
