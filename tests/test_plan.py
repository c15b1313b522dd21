"""CPU tests of the plan layer (dosecuda_amd.plan) and the NRRD files
(dosecuda_amd.nrrd): phantom set-up, plan and beam bookkeeping, output
factors, NRRD bytes and round trips, and the errors the reference's plan code
raises before any GPU work (plan.py, plan_impt.py, plan_imrt.py)."""
import gzip
import types

import numpy as np
import pytest

from dosecuda_amd import nrrd, plan as P
from dosecuda_amd.beam_models import square_field_mlc

F = np.float32


# ---------------------------------------------------------------------------
# phantom and HU handling
# ---------------------------------------------------------------------------
def test_cube_phantom_default():
    """plan.py:218-224: 138^3 voxels of 3 mm, origin at minus half the
    extent, water inside a round(10/3) = 3 voxel air margin."""
    g = P.IMPTDoseGrid()
    g.createCubePhantom()
    assert g.HU.shape == (138, 138, 138) and g.HU.dtype == F
    assert np.array_equal(g.origin, [-207.0, -207.0, -207.0])
    assert np.array_equal(g.spacing, [3.0, 3.0, 3.0])
    assert np.array_equal(g.size, [138, 138, 138])
    assert np.count_nonzero(g.HU == 0.0) == 132 ** 3
    assert np.count_nonzero(g.HU == -1000.0) == 138 ** 3 - 132 ** 3
    assert g.HU[3, 3, 3] == 0.0 and g.HU[2, 70, 70] == -1000.0 and g.HU[135, 70, 70] == -1000.0


def test_cube_phantom_margin_rounds():
    g = P.IMRTDoseGrid()
    g.createCubePhantom(size=[20, 24, 28], spacing=4.0)     # round(2.5) = 2 (banker's)
    assert g.HU.shape == (20, 24, 28)
    assert g.HU[2, 2, 2] == 0.0 and g.HU[1, 10, 10] == -1000.0
    assert np.allclose(g.origin, [-40.0, -48.0, -56.0])


def test_couch_model_override():
    """plan.py:118-123: the last int(50/s) rows of y take the HU whose
    RLSP-1 spreads couch_wet mm over that thickness."""
    g = P.DoseGrid()
    g.createCubePhantom(size=[10, 40, 12], spacing=2.5)
    g.applyCouchModel(8.0)
    assert np.all(g.HU[:, -20:, :] == F(-840.0))
    assert np.all(g.HU[:, :20, :] != F(-840.0))


def test_hu_conversions_use_tables():
    g = P.IMPTDoseGrid()
    g.createCubePhantom(size=[8, 8, 8], spacing=3.0)
    r = g.RLSPFromHU("HitachiProbeatJHU")
    assert r.dtype == F and r.shape == g.HU.shape
    assert abs(float(r[4, 4, 4]) - 1.0) < 0.05 and float(r[0, 0, 0]) < 0.01
    d = P.IMRTDoseGrid()
    d.createCubePhantom(size=[8, 8, 8], spacing=3.0)
    rho = d.DensityFromHU("VarianTrueBeamHF")
    assert abs(float(rho[4, 4, 4]) - 1.0) < 0.05 and float(rho[0, 0, 0]) < 0.01


# ---------------------------------------------------------------------------
# NRRD
# ---------------------------------------------------------------------------
def test_nrrd_header_bytes(tmp_path):
    """The header follows the NRRD0004 field grammar; sizes fastest axis
    first; data are the C-order array's bytes, little-endian."""
    a = np.arange(2 * 3 * 4, dtype=F).reshape(2, 3, 4)
    p = tmp_path / "a.nrrd"
    nrrd.write(str(p), a, (-1.5, 2.0, 3.25), (0.5, 0.5, 2.0))
    raw = p.read_bytes()
    head, data = raw.split(b"\n\n", 1)
    assert head.decode().split("\n") == [
        "NRRD0004", "type: float", "dimension: 3", "space: left-posterior-superior",
        "sizes: 4 3 2", "space directions: (0.5,0,0) (0,0.5,0) (0,0,2)",
        "kinds: domain domain domain", "endian: little", "encoding: raw",
        "space origin: (-1.5,2,3.25)"]
    assert data == a.astype("<f4").tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16, np.uint8, np.int32])
@pytest.mark.parametrize("compress", [False, True])
def test_nrrd_round_trip(tmp_path, dtype, compress):
    rng = np.random.default_rng(3)
    a = (rng.uniform(-1000, 1000, (5, 6, 7)) if np.dtype(dtype).kind == "f"
         else rng.integers(0, 100, (5, 6, 7))).astype(dtype)
    p = str(tmp_path / "v.nrrd")
    o, s = np.array([-207.0, 1.0 / 3.0, 12.5]), np.array([3.0, 2.5, 1.0 / 7.0])
    nrrd.write(p, a, o, s, compress=compress)
    b, o2, s2 = nrrd.read(p)
    assert b.dtype == a.dtype and np.array_equal(a, b)
    assert np.array_equal(o, o2) and np.array_equal(s, s2)


def test_nrrd_reads_spec_variants(tmp_path):
    """A file written by hand from the specification: comments, key/value
    pairs, a big-endian short type spelled 'signed short', ``spacings``
    instead of ``space directions``, gzip encoding and a byte skip."""
    a = np.array([[[1, -2], [300, -400]], [[5, 6], [7, -32768]]], dtype=np.int16)
    payload = b"\x00\x00\x00" + a.astype(">i2").tobytes()
    head = ("NRRD0005\n# a comment\ntype: signed short\ndimension: 3\nsizes: 2 2 2\n"
            "spacings: 1.5 2 2.5\nendian: big\nencoding: gzip\nbyte skip: 3\n"
            "modality:=CT\n\n").encode()
    p = tmp_path / "h.nrrd"
    p.write_bytes(head + gzip.compress(payload))
    b, o, s = nrrd.read(str(p))
    assert np.array_equal(a, b) and b.dtype == np.int16
    assert np.array_equal(s, [1.5, 2.0, 2.5]) and np.array_equal(o, [0.0, 0.0, 0.0])


def test_nrrd_spacing_is_direction_length(tmp_path):
    p = tmp_path / "d.nrrd"
    head = ("NRRD0004\ntype: uchar\ndimension: 3\nsizes: 1 1 2\n"
            "space directions: (0,3,4) (1,0,0) (0,0,2)\nspace origin: (1,2,3)\n"
            "encoding: raw\n\n").encode()
    p.write_bytes(head + bytes([7, 9]))
    b, o, s = nrrd.read(str(p))
    assert b.shape == (2, 1, 1) and list(b.ravel()) == [7, 9]
    assert np.array_equal(s, [5.0, 1.0, 2.0]) and np.array_equal(o, [1.0, 2.0, 3.0])


def test_nrrd_rejects(tmp_path):
    p = tmp_path / "x.nrrd"
    p.write_bytes(b"P6\n")
    with pytest.raises(ValueError, match="not an NRRD"):
        nrrd.read(str(p))
    p.write_bytes(b"NRRD0004\ntype: float\ndimension: 2\nsizes: 2 2\nencoding: raw\n\n" + bytes(16))
    with pytest.raises(ValueError, match="3-dimensional"):
        nrrd.read(str(p))
    p.write_bytes(b"NRRD0004\ntype: float\ndimension: 3\nsizes: 2 2 2\ndata file: y.raw\n\n")
    with pytest.raises(ValueError, match="detached"):
        nrrd.read(str(p))
    p.write_bytes(b"NRRD0004\ntype: float\ndimension: 3\nsizes: 2 2 2\nencoding: raw\n\n" + bytes(8))
    with pytest.raises(ValueError, match="shorter"):
        nrrd.read(str(p))
    with pytest.raises(ValueError):
        nrrd.write(str(p), np.zeros((2, 2), F), (0, 0, 0), (1, 1, 1))


def test_ct_nrrd_round_trip(tmp_path):
    g = P.DoseGrid()
    g.createCubePhantom(size=[12, 10, 14], spacing=2.5)
    p = str(tmp_path / "ct.nrrd")
    g.writeCTNRRD(p)
    h = P.DoseGrid()
    h.loadCTNRRD(p)
    assert np.array_equal(h.HU, g.HU) and h.HU.dtype == F
    assert np.array_equal(h.origin, g.origin) and np.array_equal(h.spacing, g.spacing)
    assert np.array_equal(h.size, g.size)
    with pytest.raises(Exception, match="CT path must have .nrrd extension"):
        g.writeCTNRRD(str(tmp_path / "ct.nii"))


def test_dose_nrrd_rbe_and_beams(tmp_path):
    """plan.py:165-191: EFFECTIVE multiplies by 1.1 in float32, PHYSICAL by
    1.0; individual beams go to x_beamNN.nrrd."""
    g = P.DoseGrid()
    g.createCubePhantom(size=[6, 7, 8], spacing=2.0)
    rng = np.random.default_rng(5)
    g.beam_doses = [rng.uniform(0, 2, (6, 7, 8)).astype(F) for _ in range(2)]
    g.dose = g.beam_doses[0] + g.beam_doses[1]
    p = str(tmp_path / "dose.nrrd")
    g.writeDoseNRRD(p)
    d, _, _ = nrrd.read(p)
    assert np.array_equal(d, (g.dose * 1.1).astype(F))
    g.writeDoseNRRD(p, dose_type="PHYSICAL")
    assert np.array_equal(nrrd.read(p)[0], g.dose)
    g.writeDoseNRRD(p, individual_beams=True, dose_type="PHYSICAL")
    for i in range(2):
        b, o, s = nrrd.read(str(tmp_path / ("dose_beam%02d.nrrd" % (i + 1))))
        assert np.array_equal(b, g.beam_doses[i])
        assert np.array_equal(o, g.origin) and np.array_equal(s, g.spacing)
    with pytest.raises(Exception, match="Dose path must have .nrrd extension"):
        g.writeDoseNRRD(str(tmp_path / "dose.dcm"))
    with pytest.raises(Exception, match="Unknown dose type: BIO"):
        g.writeDoseNRRD(p, dose_type="BIO")


# ---------------------------------------------------------------------------
# beams and plans
# ---------------------------------------------------------------------------
def test_impt_beam_spots():
    """plan_impt.py:261-282, including the 1-D row left by a first spot."""
    b = P.IMPTBeam()
    b.addSingleSpot(1.0, 2.0, 0.5, 7)
    assert b.spot_list.shape == (4,) and b.n_spots == 1
    b.addSingleSpot(-3.0, 4.0, 0.25, 9)
    assert b.spot_list.shape == (2, 4) and b.spot_list.dtype == F
    assert np.array_equal(b.spot_list[1], F([-3, 4, 0.25, 9]))
    cp = types.SimpleNamespace(ScanSpotPositionMap=[10.0, 11.0, 12.0, 13.0],
                               ScanSpotMetersetWeights=[0.1, 0.2])
    b.addSpotData(cp, 3)
    assert b.n_spots == 4 and b.spot_list.shape == (4, 4)
    assert np.array_equal(b.spot_list[2:], F([[10, 11, 0.1, 3], [12, 13, 0.2, 3]]))
    b.changeSpotEnergy(5)
    assert np.all(b.spot_list[:, 3] == 5)
    c = P.IMPTBeam(b)
    assert c.n_spots == 4 and c.spot_list is b.spot_list
    b.resetSpots()
    assert b.spot_list.shape == (0, 4) and b.n_spots == 0
    fresh = P.IMPTBeam()
    fresh.addSpotData(cp, 1)
    assert fresh.spot_list.shape == (2, 4)


def test_impt_plan_models():
    p = P.IMPTPlan()
    assert list(p.dicom_rangeshifter_label.astype(str)) == ["0", "4"]
    assert list(p.folder_rangeshifter_label) == ["RS0", "RS4"]
    assert len(p.beam_models) == 2
    for m in p.beam_models:
        assert m.lut_depths.shape[1] == 400 and m.lut_depths.dtype == F
        assert m.divergence_params.shape[0] == m.lut_depths.shape[0]
    assert p.beam_models[1].dicom_rangeshifter_label == "4"
    assert not np.array_equal(p.beam_models[0].lut_idds, p.beam_models[1].lut_idds)
    b = P.IMPTBeam()
    p.addBeam(b)
    p.addBeam(b)
    assert p.n_beams == 2 and p.n_fractions == 1


def test_imrt_plan_models_and_square_field():
    """plan_imrt.py:174-190, 353-385."""
    p = P.IMRTPlan("VarianTrueBeamHF")
    assert list(p.dicom_energy_label.astype(str)) == ["6", "10"]
    assert len(p.beam_models) == 2
    assert p.beam_models[0].kernel.shape[1] == 6 and p.beam_models[0].n_mlc_pairs > 0
    p.addSquareField()
    p.addSquareField('6', 15.0, 15.0, mu=50, gantry_angle=90.0, collimator_angle=20.0)
    assert p.n_beams == 2
    b0, b1 = p.beam_list
    assert b0.n_cps == 1 and b0.dicom_energy_label == '6'
    cp = b1.cp_list[0]
    assert cp.mu == 50 and cp.ga == 90.0 and cp.ca == 20.0 and cp.ta == 0.0
    assert np.array_equal(cp.mlc, square_field_mlc(p.beam_models[0], 15.0, 15.0))
    assert cp.mlc.dtype == F and cp.mlc.shape == (p.beam_models[0].n_mlc_pairs, 4)
    assert np.array_equal(cp.xjaws, F([-75, 75])) and np.array_equal(cp.yjaws, F([-75, 75]))
    assert np.array_equal(cp.iso, F([0, 0, 0]))
    e = P.IMRTPlan("ElektaVersaSibAxesse")
    assert list(e.dicom_energy_label.astype(str)) == ["6", "6FFF", "10", "15"]


def test_output_factor_square_field():
    """A square field's equivalent square 4A/P is its side, so the output
    factor is the table value at the side (plan_imrt.py:61-92)."""
    p = P.IMRTPlan()
    m = p.beam_models[0]
    p.addSquareField(dimx=10, dimy=10)
    of = m.outputFactor(p.beam_list[0].cp_list[0])
    assert of == np.interp(100.0, m.output_factor_equivalent_squares, m.output_factor_values)


def test_photon_energy_validation():
    e = P.IMRTPhotonEnergy("6")
    assert e.kernel is None and e.mu_calibration is None
    with pytest.raises(Exception, match="output_factor_equivalent_squares not set"):
        e.validate_parameters()


def test_dicom_inputs_that_are_missing_raise(tmp_path):
    """The DICOM readers (tests/test_dicom.py covers them) raise on missing
    files and on directories without an image series."""
    g = P.DoseGrid()
    for call in (lambda: g.loadCTDCM(str(tmp_path / "none")),
                 lambda: g.resampleCTfromReferenceDose(str(tmp_path / "none.dcm")),
                 lambda: P.IMPTPlan().readPlanDicom(str(tmp_path / "none.dcm")),
                 lambda: P.IMRTPlan().readPlanDicom(str(tmp_path / "none.dcm"))):
        with pytest.raises(FileNotFoundError):
            call()
    (tmp_path / "empty").mkdir()
    (tmp_path / "empty" / "notes.txt").write_text("no images here")
    with pytest.raises(ValueError, match="no DICOM image series"):
        g.loadCTDCM(str(tmp_path / "empty"))


# ---------------------------------------------------------------------------
# errors raised before any GPU work, as the reference raises them
# ---------------------------------------------------------------------------
def test_impt_plan_errors_before_gpu(capsys):
    g = P.IMPTDoseGrid()
    g.createCubePhantom(size=[8, 8, 8], spacing=3.0)
    p = P.IMPTPlan()
    b = P.IMPTBeam()
    b.dicom_rangeshifter_label = "7"
    b.addSingleSpot(0.0, 0.0, 1.0, 10)
    b.addSingleSpot(1.0, 0.0, 1.0, 10)
    p.addBeam(b)
    with pytest.raises(SystemExit) as e:
        g.computeIMPTPlan(p)
    assert e.value.code == 1
    assert "Beam model not found for rangeshifter ID 7" in capsys.readouterr().out
    # one spot leaves a 1-D spot_list, which the proton entry point rejects
    b1 = P.IMPTBeam()
    b1.dicom_rangeshifter_label = "0"
    b1.addSingleSpot(0.0, 0.0, 1.0, 10)
    p1 = P.IMPTPlan()
    p1.addBeam(b1)
    with pytest.raises(ValueError, match="'spot_list' must be 2-dimensional and of type float."):
        g.computeIMPTPlan(p1)
    g.spacing = np.array([3.0, 3.0, 2.5])
    with pytest.raises(Exception, match="Spacing must be isotropic"):
        g.computeIMPTPlan(p1)


def test_imrt_plan_errors_before_gpu(capsys):
    g = P.IMRTDoseGrid()
    g.createCubePhantom(size=[8, 8, 8], spacing=3.0)
    p = P.IMRTPlan()
    p.addSquareField('18')
    with pytest.raises(SystemExit):
        g.computeIMRTPlan(p)
    assert "Beam model not found for beam energy 18" in capsys.readouterr().out
    q = P.IMRTPlan()
    q.addSquareField()
    q.beam_list[0].cp_list[0].mu = "100"
    with pytest.raises(TypeError, match="must be real number, not str"):
        g.computeIMRTPlan(q)
    g.spacing = np.array([2.0, 3.0, 3.0])
    with pytest.raises(Exception, match="Spacing must be isotropic"):
        g.computeIMRTPlan(q)


def test_empty_plans_stay_on_host():
    g = P.IMPTDoseGrid()
    g.createCubePhantom(size=[6, 6, 6], spacing=2.0)
    p = P.IMPTPlan()
    p.n_fractions = 3
    g.computeIMPTPlan(p)
    assert g.dose.shape == (6, 6, 6) and not g.dose.any() and g.beam_doses == []
    h = P.IMRTDoseGrid()
    h.createCubePhantom(size=[6, 6, 6], spacing=2.0)
    h.computeIMRTPlan(P.IMRTPlan())
    assert not h.dose.any() and h.dose_device is None
