"""DICOM, NIfTI and CT resampling of the plan layer (CPU).

The reference reads and writes these files with pydicom and SimpleITK
(plan.py:35-163,206-216, plan_imrt.py:273-351, plan_impt.py:298-643), which
are not installed here and ship no fixture files, so the comparisons with the
reference's own outputs stay open; what is pinned:
* the reader against byte strings assembled here by hand from the encoding
  rules of PS3.5 / PS3.10 (explicit and implicit VR little endian, sequences
  and items of defined and undefined length, an unknown-VR sequence);
* the writer by reading its files back;
* loadCTDCM / resampling / writeDoseDCM / readPlanDicom / writePlanDicom
  against values computed by hand from the reference's code (slope and
  intercept, the -1000 clip, the z sort, the RT Dose scaling, the meterset
  differences, the reference's jaw / MLC item indices, the layer pairs);
* the linear resampling against closed forms (identity grid, integer
  shifts, exact reproduction of linear functions, the half-voxel border).
"""
import os
import struct

import numpy as np
import pytest

from dosecuda_amd import dicom as D
from dosecuda_amd import nifti, plan as P


# ---------------------------------------------------------------------------
# hand-assembled encodings
# ---------------------------------------------------------------------------
def _ex(g, e, vr, raw):
    """Explicit VR little endian element (PS3.5 7.1.2)."""
    if len(raw) % 2:
        raw += b" "
    if vr in ("OB", "OW", "SQ", "UN", "UT", "OF"):
        return struct.pack("<HH2sHI", g, e, vr.encode(), 0, len(raw)) + raw
    return struct.pack("<HH2sH", g, e, vr.encode(), len(raw)) + raw


def _im(g, e, raw):
    """Implicit VR little endian element (PS3.5 7.1.3)."""
    if len(raw) % 2:
        raw += b" "
    return struct.pack("<HHI", g, e, len(raw)) + raw


def _meta(ts):
    ts_raw = ts.encode() + (b"\x00" if len(ts) % 2 else b"")
    body = _ex(0x0002, 0x0002, "UI", b"1.2.840.10008.5.1.4.1.1.481.5\x00") + _ex(0x0002, 0x0010, "UI", ts_raw)
    return b"\x00" * 128 + b"DICM" + _ex(0x0002, 0x0000, "UL", struct.pack("<I", len(body))) + body


UNDEF = 0xFFFFFFFF


def test_reads_explicit_vr_with_nested_undefined_length_sequences():
    item = _ex(0x300A, 0x00C0, "IS", b"7") + _ex(0x300A, 0x00CE, "CS", b"TREATMENT")
    inner = (struct.pack("<HHI", 0xFFFE, 0xE000, UNDEF) + _ex(0x300A, 0x0112, "IS", b"0")
             + _ex(0x300A, 0x0114, "DS", b"6") + struct.pack("<HHI", 0xFFFE, 0xE00D, 0)
             + struct.pack("<HHI", 0xFFFE, 0xE000, UNDEF) + _ex(0x300A, 0x0112, "IS", b"1")
             + struct.pack("<HHI", 0xFFFE, 0xE00D, 0) + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0))
    item += struct.pack("<HH2sHI", 0x300A, 0x0111, b"SQ", 0, UNDEF) + inner
    seq = (struct.pack("<HH2sHI", 0x300A, 0x00B0, b"SQ", 0, UNDEF) + struct.pack("<HHI", 0xFFFE, 0xE000, len(item))
           + item + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0))
    body = (_ex(0x0008, 0x0060, "CS", b"RTPLAN") + _ex(0x0028, 0x0010, "US", struct.pack("<H", 512))
            + _ex(0x0028, 0x0030, "DS", b"0.976\\1.25") + _ex(0x300A, 0x0394, "FL", struct.pack("<3f", 1.5, -2.0, 3.25))
            + seq + _ex(0x7FE0, 0x0010, "OW", struct.pack("<2h", -1000, 40)))
    ds = D.read(_meta(D.EXPLICIT_LE) + body)
    assert ds.Modality == "RTPLAN" and ds.Rows == 512
    assert ds.PixelSpacing == [0.976, 1.25] and str(ds.PixelSpacing[0]) == "0.976"
    assert ds.ScanSpotPositionMap == [1.5, -2.0, 3.25]
    (b,) = ds.BeamSequence
    assert b.BeamNumber == 7 and b.TreatmentDeliveryType == "TREATMENT"
    cps = b.ControlPointSequence
    assert [c.ControlPointIndex for c in cps] == [0, 1]
    assert cps[0].NominalBeamEnergy == "6" and cps[0].NominalBeamEnergy == 6.0
    assert not hasattr(cps[1], "NominalBeamEnergy")
    assert ds.PixelData == struct.pack("<2h", -1000, 40)


def test_reads_implicit_vr_without_preamble():
    item = _im(0x300A, 0x0112, b"0") + _im(0x300A, 0x011E, b"180.5")
    seq_body = struct.pack("<HHI", 0xFFFE, 0xE000, len(item)) + item
    body = (_im(0x0008, 0x0060, b"CT") + _im(0x0020, 0x0032, b"-250\\-250\\12.5")
            + _im(0x0028, 0x0011, struct.pack("<H", 3)) + _im(0x0028, 0x1053, b"1")
            + struct.pack("<HHI", 0x300A, 0x0111, len(seq_body)) + seq_body
            + _im(0x0099, 0x0010, b"PRIVATE"))
    with pytest.raises(ValueError):
        D.read(body)
    ds = D.read(body, force=True)
    assert ds.Modality == "CT" and ds.Columns == 3 and ds.RescaleSlope == 1.0
    assert ds.ImagePositionPatient == [-250.0, -250.0, 12.5]
    assert ds.ControlPointSequence[0].GantryAngle == 180.5
    assert ds[(0x0099, 0x0010)].VR == "UN" and ds[(0x0099, 0x0010)].value == b"PRIVATE "


def test_unknown_vr_sequence_of_undefined_length_holds_implicit_items():
    item = _im(0x300A, 0x0112, b"3")
    un = (struct.pack("<HH2sHI", 0x4001, 0x1050, b"UN", 0, UNDEF) + struct.pack("<HHI", 0xFFFE, 0xE000, len(item))
          + item + struct.pack("<HHI", 0xFFFE, 0xE0DD, 0))
    ds = D.read(_meta(D.EXPLICIT_LE) + un + _ex(0x4001, 0x1051, "LO", b"after"))
    el = ds[(0x4001, 0x1050)]
    assert el.VR == "SQ" and el.value[0].ControlPointIndex == 3
    assert ds[(0x4001, 0x1051)].value == "after"


def test_compressed_transfer_syntax_is_refused():
    with pytest.raises(NotImplementedError):
        D.read(_meta("1.2.840.10008.1.2.4.70") + _ex(0x0008, 0x0060, "CS", b"CT"))


def test_ds_is_values_keep_their_text():
    labels = np.array(["6", "10", "15"])
    assert list(labels.astype(str)).index(D.DS("10")) == 1
    assert D.DS("6") != "6.0" and D.DS("6") == 6 and D.IS("12") == "12" and D.IS(" 3") == 3
    for x in (1.0 / 3.0, 123456.789012345, -1e-12, 2.5e20, 0.1):
        s = D._format_ds(x)
        assert len(s) <= 16 and abs(float(s) - x) <= 1e-10 * abs(x)


def test_write_read_round_trip(tmp_path):
    ds = D.Dataset()
    ds.SOPClassUID = D.RT_PLAN_STORAGE
    ds.SOPInstanceUID = D.generate_uid()
    ds.PatientName = "Phantom^Water"
    ds.Rows = 5
    ds.TreatmentProtocols = ["A", "B"]
    ds.IsocenterPosition = [1.5, -2.25, 100.0]
    ds.NumberOfFractionsPlanned = 30
    ds.ScanSpotMetersetWeights = [0.5, 1.25, 2.0]
    ds.SnoutPosition = 391.0
    ds.TableTopVerticalPosition = ""
    ds.PixelData = np.arange(6, dtype=np.uint16)
    cp = D.Dataset()
    cp.ControlPointIndex = 0
    cp.GantryAngle = 90
    ds.IonControlPointSequence = [cp, D.Dataset()]
    ds.add_new((0x4001, 0x0010), "LO", "CREATOR")
    ds.add_new((0x4001, 0x1002), "UN", b"abc")
    path = str(tmp_path / "x.dcm")
    D.write(ds, path)
    r = D.read(path)
    assert r.file_meta.TransferSyntaxUID == D.EXPLICIT_LE
    assert r.file_meta.MediaStorageSOPInstanceUID == ds.SOPInstanceUID
    assert r.PatientName == "Phantom^Water" and r.Rows == 5 and r.TreatmentProtocols == ["A", "B"]
    assert r.IsocenterPosition == [1.5, -2.25, 100.0] and r.NumberOfFractionsPlanned == 30
    assert r.ScanSpotMetersetWeights == [0.5, 1.25, 2.0] and r.SnoutPosition == 391.0
    assert r.TableTopVerticalPosition is None
    assert np.array_equal(np.frombuffer(r.PixelData, np.uint16), np.arange(6))
    assert r.IonControlPointSequence[0].GantryAngle == 90 and len(r.IonControlPointSequence[1]) == 0
    assert r[(0x4001, 0x0010)].value == "CREATOR" and r[(0x4001, 0x1002)].value == b"abc\x00"


# ---------------------------------------------------------------------------
# CT series, resampling, RT Dose, NIfTI
# ---------------------------------------------------------------------------
def _write_ct_series(d, hu_zyx, origin, spacing, slope=2.0, intercept=-1024.0, seed=0):
    """One CT slice per file, files in a shuffled order; stored value =
    (HU - intercept) / slope as int16."""
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(seed)
    nz, ny, nx = hu_zyx.shape
    series, frame, study = D.generate_uid(), D.generate_uid(), D.generate_uid()
    for pos, k in enumerate(rng.permutation(nz)):
        s = D.Dataset()
        s.SOPClassUID = D.CT_IMAGE_STORAGE
        s.SOPInstanceUID = D.generate_uid()
        s.Modality = "CT"
        s.PatientName = "Test^CT"
        s.PatientID = "ID42"
        s.StudyInstanceUID = study
        s.SeriesInstanceUID = series
        s.FrameOfReferenceUID = frame
        s.InstanceNumber = int(k) + 1
        s.ImagePositionPatient = [origin[0], origin[1], origin[2] + k * spacing[2]]
        s.ImageOrientationPatient = [1, 0, 0, 0, 1, 0]
        s.PixelSpacing = [spacing[1], spacing[0]]            # row spacing, column spacing
        s.SliceThickness = spacing[2]
        s.Rows, s.Columns = ny, nx
        s.SamplesPerPixel = 1
        s.PhotometricInterpretation = "MONOCHROME2"
        s.BitsAllocated, s.BitsStored, s.HighBit, s.PixelRepresentation = 16, 16, 15, 1
        s.RescaleSlope = slope
        s.RescaleIntercept = intercept
        s.PixelData = np.round((hu_zyx[k] - intercept) / slope).astype(np.int16)
        D.write(s, os.path.join(d, "img_%03d.dcm" % pos))


def test_load_ct_dcm(tmp_path):
    rng = np.random.default_rng(1)
    hu = rng.integers(-1300, 1500, size=(5, 6, 8)).astype(np.float64) * 2.0
    _write_ct_series(str(tmp_path / "ct"), hu, (-20.0, -15.5, 7.0), (1.25, 0.75, 2.5))
    g = P.IMPTDoseGrid()
    g.loadCTDCM(str(tmp_path / "ct"))
    assert g.HU.dtype == np.float32 and g.HU.shape == (5, 6, 8)
    assert np.array_equal(g.HU, np.clip(hu, -1000.0, None).astype(np.float32))
    assert np.allclose(g.origin, [-20.0, -15.5, 7.0]) and np.allclose(g.spacing, [1.25, 0.75, 2.5])
    assert np.array_equal(g.size, [5, 6, 8])
    series = P.IMPTPlan().load_ct_series(str(tmp_path / "ct"))
    assert [int(s.InstanceNumber) for s in series] == [1, 2, 3, 4, 5]


def test_resample_closed_forms():
    rng = np.random.default_rng(2)
    vol = rng.normal(size=(6, 7, 9)).astype(np.single)
    o, s = np.array([-4.0, 2.0, 10.0]), np.array([1.5, 2.0, 2.5])
    # identity grid
    same = P._resample_linear(vol, o, s, o, s, (9, 7, 6))
    assert np.array_equal(same, vol)
    # integer shift by (2, -1, 1) voxels: shifted copy, -1000 where outside
    sh = P._resample_linear(vol, o, s, o + s * np.array([2, -1, 1]), s, (9, 7, 6))
    assert np.array_equal(sh[:5, 1:, :7], vol[1:, :6, 2:])
    assert np.all(sh[5] == -1000.0) and np.all(sh[:, 0] == -1000.0) and np.all(sh[:, :, 7:] == -1000.0)
    # linear functions are reproduced exactly at interior points
    z, y, x = np.meshgrid(np.arange(6), np.arange(7), np.arange(9), indexing="ij")
    px, py, pz = o[0] + s[0] * x, o[1] + s[1] * y, o[2] + s[2] * z
    lin = (0.5 * px - 0.25 * py + 0.125 * pz + 3.0).astype(np.float64)
    no, ns, nsz = o + np.array([0.3, 0.7, 1.1]), np.array([0.9, 1.3, 1.7]), (12, 9, 6)
    r = P._resample_linear(lin, o, s, no, ns, nsz)
    zz, yy, xx = np.meshgrid(np.arange(6), np.arange(9), np.arange(12), indexing="ij")
    qx, qy, qz = no[0] + ns[0] * xx, no[1] + ns[1] * yy, no[2] + ns[2] * zz
    inner = (qx <= o[0] + s[0] * 8) & (qy <= o[1] + s[1] * 6) & (qz <= o[2] + s[2] * 5)
    expect = 0.5 * qx - 0.25 * qy + 0.125 * qz + 3.0
    assert np.allclose(r[inner], expect[inner], atol=1e-5)
    # half-voxel border: edge value inside, -1000 beyond
    b = P._resample_linear(vol, o, s, o - 0.49 * s, s, (9, 7, 6))
    assert np.array_equal(b[0, 0, 0], vol[0, 0, 0])
    b = P._resample_linear(vol, o, s, o - 0.51 * s, s, (9, 7, 6))
    assert b[0, 0, 0] == -1000.0


def test_resample_ct_from_spacing_sizes():
    g = P.IMRTDoseGrid()
    g.createCubePhantom(size=[20, 24, 30], spacing=3.0)
    g.resampleCTfromSpacing(2.0)
    assert g.spacing == (2.0, 2.0, 2.0)
    assert g.HU.shape == (int(20 * 3 / 2), int(24 * 3 / 2), int(30 * 3 / 2))
    assert g.HU[15, 18, 22] == 0.0 and g.HU[0, 0, 0] == -1000.0


def _ref_dose_file(path, shape, origin, spacing):
    ds = D.Dataset()
    ds.SOPClassUID = D.RT_DOSE_STORAGE
    ds.SOPInstanceUID = D.generate_uid()
    ds.Modality = "RTDOSE"
    ds.SeriesDescription = "TPS dose"
    ds.Rows, ds.Columns = shape[1], shape[2]
    ds.NumberOfFrames = shape[0]
    ds.SamplesPerPixel = 1
    ds.BitsAllocated, ds.BitsStored, ds.HighBit, ds.PixelRepresentation = 16, 16, 15, 0
    ds.PixelSpacing = [spacing[0], spacing[1]]
    ds.ImagePositionPatient = list(origin)
    ds.GridFrameOffsetVector = [k * spacing[2] for k in range(shape[0])]
    ds.DoseUnits = "GY"
    ds.DoseType = "PHYSICAL"
    ds.DoseSummationType = "PLAN"
    ds.DoseGridScaling = 1e-3
    rp = D.Dataset()
    rp.ReferencedSOPClassUID = D.RT_PLAN_STORAGE
    rp.ReferencedSOPInstanceUID = "1.2.3"
    ds.ReferencedRTPlanSequence = [rp]
    ds.PixelData = np.zeros(shape, np.uint16)
    D.write(ds, path)


def test_write_dose_dcm_and_resample_from_reference_dose(tmp_path):
    ref = str(tmp_path / "ref.dcm")
    _ref_dose_file(ref, (4, 5, 6), (-7.5, -10.0, -4.0), (2.5, 2.5, 2.5))
    g = P.IMPTDoseGrid()
    g.createCubePhantom(size=[12, 12, 12], spacing=2.0)
    g.resampleCTfromReferenceDose(ref)
    assert g.HU.shape == (4, 5, 6) and np.allclose(g.spacing, [2.5, 2.5, 2.5])
    assert np.allclose(g.origin, [-7.5, -10.0, -4.0])
    rng = np.random.default_rng(3)
    g.dose = rng.uniform(0, 70, size=(4, 5, 6)).astype(np.single)
    g.beam_doses = [g.dose * 0.25, g.dose * 0.75]
    with pytest.raises(Exception):
        g.writeDoseDCM(str(tmp_path / "d.nrrd"), ref)
    with pytest.raises(Exception):
        g.writeDoseDCM(str(tmp_path / "d.dcm"), ref, dose_type="BIOLOGICAL")
    g.writeDoseDCM(str(tmp_path / "d.dcm"), ref)
    r = D.read(str(tmp_path / "d.dcm"))
    assert r.DoseSummationType == "PLAN" and r.DoseType == "EFFECTIVE"
    assert r.SeriesDescription == "TPS dose_DoseCUDA" and r.DoseGridScaling == 0.002
    assert r.ReferencedRTPlanSequence[0].ReferencedSOPInstanceUID != "1.2.3"
    assert np.array_equal(r.pixel_array, np.array(g.dose * 500.0 * 1.1, dtype=np.uint16))
    g.writeDoseDCM(str(tmp_path / "d.dcm"), ref, dose_type="PHYSICAL", individual_beams=True)
    for i in range(2):
        r = D.read(str(tmp_path / ("d_beam%02i.dcm" % (i + 1))))
        assert r.DoseSummationType == "BEAM"
        assert np.array_equal(r.pixel_array, np.array(g.beam_doses[i] * 500.0, dtype=np.uint16))


def test_write_ct_nifti(tmp_path):
    g = P.IMRTDoseGrid()
    g.createCubePhantom(size=[6, 7, 8], spacing=2.5)
    with pytest.raises(Exception):
        g.writeCTNIFTI(str(tmp_path / "ct.nii"))
    g.writeCTNIFTI(str(tmp_path / "ct.nii.gz"))
    a, o, s = nifti.read(str(tmp_path / "ct.nii.gz"))
    assert np.array_equal(a, g.HU) and np.allclose(o, g.origin) and np.allclose(s, g.spacing)
    import gzip
    raw = gzip.open(str(tmp_path / "ct.nii.gz")).read()
    srow_x = struct.unpack_from("<4f", raw, 280)
    assert srow_x[0] == -2.5 and srow_x[3] == pytest.approx(-g.origin[0])   # LPS -> RAS


# ---------------------------------------------------------------------------
# RT plans
# ---------------------------------------------------------------------------
def _bldp(types_positions):
    out = []
    for t, pos in types_positions:
        d = D.Dataset()
        d.RTBeamLimitingDeviceType = t
        d.LeafJawPositions = list(pos)
        out.append(d)
    return out


def _cp(index, weight, **kw):
    c = D.Dataset()
    c.ControlPointIndex = index
    c.CumulativeMetersetWeight = weight
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_read_imrt_plan_dicom(tmp_path):
    pl = P.IMRTPlan()
    bm = pl.beam_models[0]
    n = int(bm.n_mlc_pairs)
    leaves_a = np.concatenate([np.full(n, -30.0), np.full(n, 40.0)])
    leaves_b = np.concatenate([np.full(n, -10.0), np.full(n, 20.0)])
    beam = D.Dataset()
    beam.BeamNumber = 1
    beam.TreatmentDeliveryType = "TREATMENT"
    beam.ControlPointSequence = [
        _cp(0, 0.0, NominalBeamEnergy="6", IsocenterPosition=[10.0, -5.0, 2.5], GantryAngle=180.0,
            BeamLimitingDeviceAngle=15.0, PatientSupportAngle=350.0,
            BeamLimitingDevicePositionSequence=_bldp([("ASYMX", (-50, 50)), ("ASYMY", (-60, 60)),
                                                      ("MLCX", leaves_a)])),
        _cp(1, 0.4, GantryAngle=182.0,
            BeamLimitingDevicePositionSequence=_bldp([("ASYMX", (-50, 50)), ("ASYMY", (-60, 60)),
                                                      ("MLCX", leaves_b)])),
        _cp(2, 1.0),
    ]
    setup = D.Dataset()
    setup.TreatmentDeliveryType = "SETUP"
    jaws = D.Dataset()
    jaws.BeamNumber = 3
    jaws.TreatmentDeliveryType = "TREATMENT"
    jaws.ControlPointSequence = [
        _cp(0, 0.0, NominalBeamEnergy="10", IsocenterPosition=[0.0, 0.0, 0.0], GantryAngle=90.0,
            BeamLimitingDeviceAngle=0.0, PatientSupportAngle=0.0,
            BeamLimitingDevicePositionSequence=_bldp([("X", (-35, 45)), ("Y", (-20, 30))])),
        _cp(1, 1.0),
    ]
    fg = D.Dataset()
    fg.NumberOfFractionsPlanned = 25
    rb = D.Dataset()
    rb.ReferencedBeamNumber = 1
    rb.BeamMeterset = 200.0
    fg.ReferencedBeamSequence = [rb]
    ds = D.Dataset()
    ds.SOPClassUID = D.RT_PLAN_STORAGE
    ds.SOPInstanceUID = D.generate_uid()
    ds.FractionGroupSequence = [fg]
    ds.BeamSequence = [beam, setup, jaws]
    path = str(tmp_path / "plan.dcm")
    D.write(ds, path)

    pl.readPlanDicom(path)
    assert pl.n_fractions == 25.0 and pl.n_beams == 2
    b0, b1 = pl.beam_list
    assert list(pl.dicom_energy_label.astype(str)).index(b0.dicom_energy_label) >= 0
    assert np.array_equal(b0.iso, np.array([10.0, -5.0, 2.5], np.single))
    assert b0.n_cps == 2
    c0, c1 = b0.cp_list
    assert c0.mu == pytest.approx(0.4 * 200.0) and c1.mu == pytest.approx(0.6 * 200.0)
    assert (c0.ga, c0.ca, c0.ta) == (180.0, 15.0, 350.0) and c1.ga == 182.0
    for c, leaves in ((c0, leaves_a), (c1, leaves_b)):
        assert c.mlc.shape == (n, 4) and c.mlc.dtype == np.single
        assert np.array_equal(c.mlc[:, 0], leaves[:n]) and np.array_equal(c.mlc[:, 1], leaves[n:])
        assert np.array_equal(c.mlc[:, 2], bm.mlc_offsets.astype(np.single))
        assert np.array_equal(c.mlc[:, 3], bm.mlc_widths.astype(np.single))
    # jaw-only beam: the reference takes both jaws from device item 1
    (cj,) = b1.cp_list
    assert cj.mu == pytest.approx(200.0)
    y = np.array([-20.0, 30.0], np.single)
    closed = (bm.mlc_offsets < y[0]) | (bm.mlc_offsets > y[1])
    assert np.array_equal(cj.xjaws, y) and np.array_equal(cj.yjaws, y)
    assert np.all(cj.mlc[closed, :2] == 0.0)
    assert np.all(cj.mlc[~closed, 0] == y[0]) and np.all(cj.mlc[~closed, 1] == y[1])


def _struct_file(path, names):
    rs = D.Dataset()
    rs.SOPClassUID = D.RT_STRUCT_STORAGE
    rs.SOPInstanceUID = D.generate_uid()
    rois = []
    for i, nm in enumerate(names):
        r = D.Dataset()
        r.ROINumber = i + 1
        r.ROIName = nm
        rois.append(r)
    rs.StructureSetROISequence = rois
    D.write(rs, path)
    return rs.SOPInstanceUID


def test_impt_plan_dicom_round_trip(tmp_path):
    ct = str(tmp_path / "ct")
    _write_ct_series(ct, np.zeros((3, 4, 5)), (0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    pl = P.IMPTPlan()
    pl.n_fractions = 5
    rng = np.random.default_rng(4)
    for label, gantry, couch in (("0", 0.0, 0.0), ("4", 270.0, 10.0)):
        b = P.IMPTBeam()
        b.dicom_rangeshifter_label = label
        b.gantry_angle, b.couch_angle = gantry, couch
        b.iso = np.array([1.0, -2.0, 3.5], np.single)
        for e in (12, 3, 12, 40, 3, 3):
            b.addSingleSpot(float(rng.uniform(-50, 50)), float(rng.uniform(-50, 50)),
                            float(rng.uniform(0.1, 2.0)), e)
        pl.addBeam(b)
    uid = _struct_file(str(tmp_path / "rs.dcm"), ["BODY", "PTV"])
    out = str(tmp_path / "ion.dcm")
    pl.writePlanDicom(ct, out, str(tmp_path / "rs.dcm"))
    ds = D.read(out)
    assert ds.Modality == "RTPLAN" and ds.PatientID == "ID42"
    assert ds.ReferencedStructureSetSequence[0].ReferencedSOPInstanceUID == uid
    assert ds.DoseReferenceSequence[0].ReferencedROINumber == 2
    ib = ds.IonBeamSequence[1]
    assert ib.NumberOfControlPoints == 6 and ib.RangeShifterSequence[0].RangeShifterID == "4"
    cps = ib.IonControlPointSequence
    assert [c.ControlPointIndex for c in cps] == list(range(6))
    w = [c.CumulativeMetersetWeight for c in cps]
    assert w[0] == 0.0 and all(w[i] <= w[i + 1] for i in range(5))
    assert cps[1].ScanSpotMetersetWeights == [0.0] * len(cps[1].ScanSpotMetersetWeights) or \
        cps[1].ScanSpotMetersetWeights == 0.0

    back = P.IMPTPlan()
    back.readPlanDicom(out)
    assert back.n_fractions == 5.0 and back.n_beams == 2
    for a, b in zip(pl.beam_list, back.beam_list):
        assert b.dicom_rangeshifter_label == a.dicom_rangeshifter_label
        assert (b.gantry_angle, b.couch_angle) == (a.gantry_angle, a.couch_angle)
        assert np.array_equal(b.iso, a.iso)
        # spots come back grouped by layer in ascending energy, in their order
        order = np.argsort(a.spot_list[:, 3], kind="stable")
        exp = a.spot_list[order]
        got = b.spot_list
        assert got.shape == exp.shape
        assert np.array_equal(got[:, :3], exp[:, :3])
        m = pl.beam_models[0]
        assert np.array_equal(m.energy_labels[got[:, 3].astype(int)], m.energy_labels[exp[:, 3].astype(int)])

    _struct_file(str(tmp_path / "rs2.dcm"), ["BODY"])
    with pytest.raises(ValueError):
        pl.writePlanDicom(ct, out, str(tmp_path / "rs2.dcm"))
