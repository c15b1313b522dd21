"""Plan-level layer: dose grids, beams and plans with the class and method
names of the reference's Python API, computing on the device engine.

Mirrors DoseCUDA/plan.py, plan_impt.py and plan_imrt.py:

  DoseGrid                 plan.py:14-24          createCubePhantom plan.py:218-224
  DoseGrid.loadCTNRRD      plan.py:25-33          applyCouchModel   plan.py:118-123
  DoseGrid.writeDoseNRRD   plan.py:165-191        writeCTNRRD       plan.py:193-204
  Beam, Plan               plan.py:227-245
  IMPTBeamModel            plan_impt.py:17-104    IMPTBeam          plan_impt.py:244-282
  IMPTDoseGrid             plan_impt.py:106-242   IMPTPlan          plan_impt.py:284-296
  IMRTPhotonEnergy         plan_imrt.py:13-92     IMRTControlPoint  plan_imrt.py:95-105
  IMRTDoseGrid             plan_imrt.py:108-158   IMRTBeam          plan_imrt.py:161-171
  IMRTPlan                 plan_imrt.py:174-271, addSquareField plan_imrt.py:353-385

What changes is where the arrays live.  The reference passes host arrays to
``dose_kernels`` once per field or control point, so every call uploads the
whole volume and downloads a whole dose, and sums the doses with numpy.
Here a plan uploads the RLSP or density volume once, keeps each beam's WET,
dose and the plan sum in HBM, and downloads the per-beam and plan doses once
at the end.  The sums are done in float32 in the reference's order with the
reference's roundings (numpy < 2 computes ``cp_dose * output_factor`` in
float32: the float64 factor is rounded to float32 first), so every dose
array equals the reference loop's bit for bit.  The device doses stay
available as ``dose_device`` and ``beam_doses_device`` (torch tensors) for
callers that go on computing on the GPU.

Inputs are read with the drop-in's checks (dose_kernels' attribute rules), so
a malformed beam or model raises what ``dose_kernels`` raises.

DICOM files (CT series, RT Dose, RT Structure Set, RT Plan, RT Ion Plan) are
read and written by ``dosecuda_amd.dicom`` (uncompressed little-endian
transfer syntaxes), NRRD by ``dosecuda_amd.nrrd`` and NIfTI by
``dosecuda_amd.nifti``; the CT resampling methods interpolate linearly like
SimpleITK's ResampleImageFilter (half-voxel border, -1000 HU outside).
"""
from __future__ import annotations

import os
import sys
from struct import error as struct_error

import numpy as np

from . import beam_models as _bm
from . import dicom as _dcm
from . import dose_kernels
from . import nifti as _nifti
from . import nrrd as _nrrd
from .beam_models import VolumeObject, aperture_rows

__all__ = ["VolumeObject", "DoseGrid", "Beam", "Plan", "IMPTBeamModel", "IMPTDoseGrid",
           "IMPTBeam", "IMPTPlan", "IMRTPhotonEnergy", "IMRTControlPoint", "IMRTDoseGrid",
           "IMRTBeam", "IMRTPlan"]


def _axis_taps(n_in, o_in, s_in, n_out, o_out, s_out):
    """Per output index along one axis: the two input indices, the weight
    of the second and whether the point is inside the input.  The continuous
    index c = (o_out + s_out i - o_in) / s_in is inside for -0.5 <= c <
    n_in - 0.5; within the half-voxel border the neighbours clamp to the edge
    (ITK's linear interpolator and ResampleImageFilter)."""
    c = (float(o_out) + float(s_out) * np.arange(int(n_out), dtype=np.float64) - float(o_in)) / float(s_in)
    inside = (c >= -0.5) & (c < n_in - 0.5)
    f = np.floor(c)
    w = c - f
    i0 = np.clip(f, 0, n_in - 1).astype(np.int64)
    i1 = np.clip(f + 1, 0, n_in - 1).astype(np.int64)
    return i0, i1, w, inside


def _resample_linear(vol, origin, spacing, new_origin, new_spacing, new_size_xyz, default=-1000.0):
    """Trilinear resampling of a (z, y, x) volume with identity direction
    onto the grid new_origin + new_spacing * index, index < new_size (x, y,
    z order, as SimpleITK takes them); points outside the input take
    ``default``.  Separable: x, then y, then z, in float64."""
    v = np.asarray(vol, dtype=np.float64)
    nz, ny, nx = v.shape
    tx = _axis_taps(nx, origin[0], spacing[0], new_size_xyz[0], new_origin[0], new_spacing[0])
    ty = _axis_taps(ny, origin[1], spacing[1], new_size_xyz[1], new_origin[1], new_spacing[1])
    tz = _axis_taps(nz, origin[2], spacing[2], new_size_xyz[2], new_origin[2], new_spacing[2])
    a = v[:, :, tx[0]] * (1.0 - tx[2]) + v[:, :, tx[1]] * tx[2]
    a = a[:, ty[0], :] * (1.0 - ty[2])[None, :, None] + a[:, ty[1], :] * ty[2][None, :, None]
    a = a[tz[0], :, :] * (1.0 - tz[2])[:, None, None] + a[tz[1], :, :] * tz[2][:, None, None]
    inside = tz[3][:, None, None] & ty[3][None, :, None] & tx[3][None, None, :]
    return np.where(inside, a, default).astype(np.single)


def _ct_series(ct_dir):
    """The image slices of the first series in ct_dir (files in name
    order; the series of the first file that holds pixel data)."""
    slices = []
    series = None
    for f in _dcm.list_files(ct_dir):
        try:
            ds = _dcm.read(f, force=True)
        except (ValueError, NotImplementedError, struct_error):
            continue
        if "PixelData" not in ds or "ImagePositionPatient" not in ds:
            continue
        uid = ds.get("SeriesInstanceUID")
        if series is None:
            series = uid
        if uid == series:
            slices.append(ds)
    if not slices:
        raise ValueError("no DICOM image series found in %s" % ct_dir)
    return slices


def _torch_device(gpu_id):
    """torch and the device of ``gpu_id``; errors read like the engine's
    (RuntimeError "CUDA error: ..."), as the reference's kernels raise them."""
    import torch
    gpu = dose_kernels._gpu(gpu_id)
    if not torch.cuda.is_available():
        raise RuntimeError("CUDA error: no GPU is visible (the plan layer computes on the GPU)")
    if gpu < 0 or gpu >= torch.cuda.device_count():
        raise RuntimeError("CUDA error: invalid device ordinal")
    return torch, torch.device("cuda", gpu)


def _volume_object(data, origin, spacing):
    v = VolumeObject()
    v.voxel_data = np.array(data, dtype=np.single)
    v.origin = np.array(origin, dtype=np.single)
    v.spacing = np.array(spacing, dtype=np.single)
    return v


class _Resident:
    """A plan's volume and dose sum in HBM, uploaded at the first kernel
    call, so that the input errors the reference raises before any GPU work
    are raised here before it too.  With ``table`` = (HU nodes, values) the
    data is the HU volume: it is uploaded and converted on the device
    (float32(np.interp) bit for bit), which replaces the host conversion."""

    def __init__(self, data, gpu_id, table=None):
        self.data, self.gpu_id, self.table, self.torch = data, gpu_id, table, None

    def up(self):
        if self.torch is None:
            self.torch, self.dev = _torch_device(self.gpu_id)
            # torch's elementwise work on these tensors runs on this stream
            # too, so the engine's launches and the sums stay in order
            self.stream = self.torch.cuda.current_stream(self.dev)
            if self.table is None:
                self.vol = self.torch.from_numpy(np.ascontiguousarray(self.data)).to(self.dev)
            else:
                from . import engine
                hu = np.asarray(self.data)
                if hu.dtype not in (np.float32, np.float64):
                    hu = hu.astype(np.float64)          # np.interp computes in float64
                hu_t = self.torch.from_numpy(np.ascontiguousarray(hu)).to(self.dev)
                self.vol = engine.hu_convert(hu_t, self.table[0], self.table[1], stream=self.stream)
                del hu_t
            self.sum = self.torch.zeros_like(self.vol)
        return self

    def zeros(self):
        return self.up().torch.zeros_like(self.vol)

    def host(self):
        return self.up().vol.cpu().numpy()


def _checks_volume(hu, origin, spacing):
    """A VolumeObject for the drop-in's argument checks before the converted
    volume exists: voxel_data is a zero-stride float32 view with the HU
    volume's shape, which the checks (rank, dtype) treat exactly as the
    converted array."""
    v = VolumeObject()
    v.voxel_data = np.broadcast_to(np.float32(0.0), np.shape(hu))
    v.origin = np.array(origin, dtype=np.single)
    v.spacing = np.array(spacing, dtype=np.single)
    return v


def _model_index(labels, label, what):
    """The reference's lookup: print and exit when no model has the label
    (plan_impt.py:141-146, plan_imrt.py:145-150)."""
    try:
        return list(np.asarray(labels).astype(str)).index(label)
    except ValueError:
        print("Beam model not found for %s %s" % (what, label))
        sys.exit(1)


# ---------------------------------------------------------------------------
# grid, beam and plan bases (plan.py)
# ---------------------------------------------------------------------------
class DoseGrid:

    def __init__(self):
        self.origin = np.array([0.0, 0.0, 0.0], dtype=np.single)
        self.spacing = np.array([0.0, 0.0, 0.0], dtype=np.single)
        self.size = np.array([0, 0, 0])
        self.HU = []
        self.dose = []
        self.beam_doses = []
        self.FrameOfReferenceUID = ""
        self.dose_device = None
        self.beam_doses_device = []

    # --- CT in -------------------------------------------------------------
    def loadCTNRRD(self, ct_path):
        hu, origin, spacing = _nrrd.read(ct_path)
        self.origin = origin
        self.spacing = spacing
        self.HU = np.array(hu, dtype=np.single)
        self.size = np.array(self.HU.shape)

    def loadCTDCM(self, ct_path):
        """A CT series directory (plan.py:35-49): slices sorted by
        ImagePositionPatient z, HU = stored value * RescaleSlope +
        RescaleIntercept clipped below at -1000; origin = the first slice's
        position, spacing = (column spacing, row spacing, distance between
        the first two slices)."""
        slices = _ct_series(ct_path)
        slices.sort(key=lambda d: float(d.ImagePositionPatient[2]))
        hu = np.empty((len(slices), int(slices[0].Rows), int(slices[0].Columns)), dtype=np.float64)
        for k, d in enumerate(slices):
            px = d.pixel_array
            if px.shape != hu.shape[1:]:
                raise ValueError("CT slices of different sizes in %s" % ct_path)
            hu[k] = px * float(d.get("RescaleSlope", 1.0) or 1.0) + float(d.get("RescaleIntercept", 0.0) or 0.0)
        ps = slices[0].PixelSpacing
        if len(slices) > 1:
            dz = abs(float(slices[1].ImagePositionPatient[2]) - float(slices[0].ImagePositionPatient[2]))
        else:
            dz = float(slices[0].get("SliceThickness", 1.0) or 1.0)
        self.origin = np.array([float(v) for v in slices[0].ImagePositionPatient])
        self.spacing = np.array([float(ps[1]), float(ps[0]), dz])
        self.HU = np.clip(np.array(hu, dtype=np.single), -1000.0, None)
        self.size = np.array(self.HU.shape)

    def resampleCT(self, new_spacing, new_size, new_origin):
        """Linear resampling onto origin/spacing/size (x, y, z), -1000 HU
        outside the CT (plan.py:51-67)."""
        self.HU = _resample_linear(self.HU, self.origin, self.spacing, new_origin, new_spacing, new_size)
        self.origin = new_origin
        self.spacing = new_spacing
        self.size = np.array(self.HU.shape)

    def resampleCTfromSpacing(self, spacing):
        """Isotropic resampling from the same origin; each axis keeps
        int(extent / spacing) voxels (plan.py:69-89)."""
        sp_new = (spacing, spacing, spacing)
        sz_new = (int(self.size[2] * self.spacing[0] / sp_new[0]),
                  int(self.size[1] * self.spacing[1] / sp_new[1]),
                  int(self.size[0] * self.spacing[2] / sp_new[2]))
        self.HU = _resample_linear(self.HU, self.origin, self.spacing, self.origin, sp_new, sz_new)
        self.spacing = sp_new
        self.size = np.array(self.HU.shape)

    def resampleCTfromReferenceDose(self, ref_dose_path):
        """Resampling onto an RT Dose grid (plan.py:91-116): spacing
        (PixelSpacing[0], PixelSpacing[1], frame offset step), origin its
        ImagePositionPatient, size its pixel array's."""
        ref_dose = _dcm.read(ref_dose_path, force=True)
        gfov = ref_dose.GridFrameOffsetVector
        slice_thickness = float(gfov[1]) - float(gfov[0])
        ref_spacing = np.array([float(ref_dose.PixelSpacing[0]), float(ref_dose.PixelSpacing[1]),
                                slice_thickness])
        ref_origin = np.array([float(v) for v in ref_dose.ImagePositionPatient])
        shape = ref_dose.pixel_array.shape
        if len(shape) == 2:
            shape = (1,) + shape
        self.HU = _resample_linear(self.HU, self.origin, self.spacing, ref_origin, ref_spacing,
                                   (shape[2], shape[1], shape[0]))
        self.size = np.array(self.HU.shape)
        self.origin = ref_origin
        self.spacing = ref_spacing

    def applyCouchModel(self, couch_wet=8.0):
        """HU override of the bottom 5 cm of rows so that they hold the
        couch's water-equivalent thickness (plan.py:118-123)."""
        spacing = self.spacing[0]
        n_voxels = int(50.0 / spacing)
        hu_override_value = ((couch_wet / (n_voxels * spacing)) - 1.0) * 1000.0
        self.HU[:, -n_voxels:, :] = hu_override_value

    def createCubePhantom(self, size=[138, 138, 138], spacing=3.0):
        """Water cube with a 1 cm air margin (plan.py:218-224)."""
        self.origin = np.array([-size[0] * spacing / 2.0, -size[1] * spacing / 2.0,
                                -size[2] * spacing / 2.0])
        self.spacing = np.array([spacing, spacing, spacing])
        self.size = np.array(size)
        edge = round(10.0 / spacing)
        self.HU = np.ones(size, dtype=np.single) * -1000.0
        self.HU[edge:-edge, edge:-edge, edge:-edge] = 0.0

    # --- exports -----------------------------------------------------------
    def writeDoseDCM(self, dose_path, ref_dose_path, dose_type="EFFECTIVE", individual_beams=False):
        """The plan (or each beam's) dose times the RBE into a copy of an
        RT Dose file: unsigned 16-bit at 500 per Gy, DoseGridScaling 0.002, a
        new referenced plan UID (plan.py:125-163)."""
        if not dose_path.endswith(".dcm"):
            raise Exception("Dose path must have .dcm extension")
        if dose_type == "EFFECTIVE":
            RBE = 1.1
        elif dose_type == "PHYSICAL":
            RBE = 1.0
        else:
            raise Exception("Unknown dose type: %s" % dose_type)

        def one(dose, summation, path):
            ref_dose = _dcm.read(ref_dose_path)
            ref_dose.SeriesDescription = ref_dose.SeriesDescription + "_DoseCUDA"
            ref_dose.DoseSummationType = summation
            ref_dose.DoseType = dose_type
            ref_dose.ReferencedRTPlanSequence[0].ReferencedSOPInstanceUID = _dcm.generate_uid()
            ref_dose.PixelData = np.array(dose * 500.0 * RBE, dtype=np.uint16)
            ref_dose.DoseGridScaling = 0.002
            _dcm.write(ref_dose, path)

        if individual_beams:
            for i, beam_dose in enumerate(self.beam_doses):
                one(beam_dose, "BEAM", dose_path.replace(".dcm", "_beam%02i.dcm" % (i + 1)))
        else:
            one(self.dose, "PLAN", dose_path)

    def writeDoseNRRD(self, dose_path, individual_beams=False, dose_type="EFFECTIVE"):
        """Plan dose (or one file per beam) times the RBE: 1.1 for
        EFFECTIVE, 1.0 for PHYSICAL (plan.py:165-191)."""
        if not dose_path.endswith(".nrrd"):
            raise Exception("Dose path must have .nrrd extension")
        if dose_type == "EFFECTIVE":
            RBE = 1.1
        elif dose_type == "PHYSICAL":
            RBE = 1.0
        else:
            raise Exception("Unknown dose type: %s" % dose_type)
        if individual_beams:
            for i, beam_dose in enumerate(self.beam_doses):
                _nrrd.write(_nrrd.beam_path(dose_path, i),
                            np.array(beam_dose * RBE, dtype=np.single), self.origin, self.spacing)
        else:
            _nrrd.write(dose_path, np.array(self.dose * RBE, dtype=np.single), self.origin,
                        self.spacing)

    def writeCTNRRD(self, ct_path):
        if not ct_path.endswith(".nrrd"):
            raise Exception("CT path must have .nrrd extension")
        _nrrd.write(ct_path, np.asarray(self.HU), self.origin, self.spacing)

    def writeCTNIFTI(self, ct_path):
        if not ct_path.endswith(".nii.gz"):
            raise Exception("CT path must have .nii.gz extension")
        _nifti.write(ct_path, np.asarray(self.HU), self.origin, self.spacing)

    # --- shared device plumbing --------------------------------------------
    def _check_isotropic(self):
        if self.spacing[0] != self.spacing[1] or self.spacing[0] != self.spacing[2]:
            raise Exception("Spacing must be isotropic for IMPT dose calculation - consider "
                            "resampling CT")

    def _finish(self, res, beam_ts, n_fractions, scale_beams):
        """Scale by the fraction count and download once.  ``scale_beams``
        follows the reference: IMPT keeps fraction-scaled beam doses
        (plan_impt.py:157), IMRT unscaled ones (plan_imrt.py:155).  A plan
        without beams never touched the GPU and keeps the host zeros."""
        if res.torch is None:
            self.dose *= n_fractions
            self.dose_device, self.beam_doses_device = None, []
            return
        beams = [b * n_fractions for b in beam_ts] if scale_beams else list(beam_ts)
        res.sum.mul_(n_fractions)
        self.dose_device = res.sum
        self.beam_doses_device = beams
        self.dose = res.sum.cpu().numpy()
        self.beam_doses = [b.cpu().numpy() for b in beams]


class Beam:

    def __init__(self):
        self.iso = np.array([0.0, 0.0, 0.0], dtype=np.single)
        self.gantry_angle = 0.0
        self.collimator_angle = 0.0
        self.couch_angle = 0.0


class Plan:

    def __init__(self):
        self.n_beams = 0
        self.n_fractions = 1
        self.beam_list = []

    def addBeam(self, beam):
        self.beam_list.append(beam)
        self.n_beams += 1


# ---------------------------------------------------------------------------
# protons (plan_impt.py)
# ---------------------------------------------------------------------------
class IMPTBeamModel(_bm.ProtonBeamModel):
    """One range-shifter model of a machine, with the reference's constructor
    arguments: ``path_to_model`` is ``lookuptables/protons/<machine>``."""

    def __init__(self, dicom_rangeshifter_label, path_to_model, folder_rangeshifter_label):
        machine = os.path.basename(os.path.normpath(path_to_model))
        super().__init__(machine, folder_rangeshifter_label, dicom_rangeshifter_label)


class IMPTBeam(Beam):

    def __init__(self, beam=None):
        super().__init__()
        self.spot_list = []
        self.n_spots = 0
        self.dicom_rangeshifter_label = None
        if beam is not None:
            self.gantry_angle = beam.gantry_angle
            self.couch_angle = beam.couch_angle
            self.iso = beam.iso
            self.spot_list = beam.spot_list
            self.n_spots = beam.n_spots
            self.dicom_rangeshifter_label = beam.dicom_rangeshifter_label

    def addSpotData(self, cp, energy_id):
        """Append the spots of one ion control point (an object with
        ScanSpotPositionMap and ScanSpotMetersetWeights)."""
        spm = np.reshape(np.array(cp.ScanSpotPositionMap), (-1, 2))
        mus = np.array(cp.ScanSpotMetersetWeights)
        energy_id_array = np.full(mus.size, energy_id)
        spot_list = np.array(np.column_stack((spm, mus, energy_id_array)), dtype=np.single)
        self.spot_list = np.vstack((self.spot_list, spot_list)) if self.n_spots > 0 else spot_list
        self.n_spots += mus.size

    def changeSpotEnergy(self, energy_id):
        self.spot_list[:, 3] = energy_id

    def addSingleSpot(self, x, y, mu, energy_id):
        # the first spot is stored as a 1-D row, as in the reference; the
        # engine then rejects the field until a second spot makes it 2-D
        spot = np.array([x, y, mu, energy_id], dtype=np.single)
        self.spot_list = np.vstack((self.spot_list, spot)) if self.n_spots > 0 else spot
        self.n_spots += 1

    def resetSpots(self):
        self.spot_list = np.empty((0, 4), dtype=np.single)
        self.n_spots = 0


class IMPTPlan(Plan):

    def __init__(self, machine_name="HitachiProbeatJHU"):
        super().__init__()
        self.machine_name = machine_name
        t = _bm._read_table(os.path.join(_bm.TABLES, "protons", machine_name,
                                         "rangeshifter_labels.csv"))
        self.dicom_rangeshifter_label = np.array(t["dicom_rangeshifter_label"])
        self.folder_rangeshifter_label = np.array(t["folder_rangeshifter_label"])
        self.beam_models = [IMPTBeamModel(d, os.path.join("lookuptables", "protons", machine_name), f)
                            for d, f in zip(self.dicom_rangeshifter_label,
                                            self.folder_rangeshifter_label)]

    def load_ct_series(self, ct_dir):
        """All .dcm files of ct_dir, read and sorted by InstanceNumber
        (plan_impt.py:298-305)."""
        files = [os.path.join(ct_dir, f) for f in os.listdir(ct_dir) if f.lower().endswith(".dcm")]
        dsets = [_dcm.read(f) for f in files]
        dsets.sort(key=lambda ds: ds.InstanceNumber)
        return dsets

    def writePlanDicom(self, ct_dir, output_file, rt_struct_file=None):
        """An RT Ion Plan for the plan's pencil-beam fields (plan_impt.py:307-603):
        patient, study and frame of reference from the CT series, one fraction
        group, and per field a pair of ion control points per energy layer (the
        layer's spots and weights, then the same spots with zero weight)."""
        import datetime
        ct_series = self.load_ct_series(ct_dir)
        first = ct_series[0]
        ds = _dcm.Dataset()
        sop_uid = _dcm.generate_uid()
        ds.PatientName = first.PatientName
        ds.PatientID = first.PatientID
        ds.StudyInstanceUID = first.StudyInstanceUID
        ds.SeriesInstanceUID = _dcm.generate_uid()
        ds.FrameOfReferenceUID = first.FrameOfReferenceUID
        ds.SOPClassUID = _dcm.RT_ION_PLAN_STORAGE
        ds.SOPInstanceUID = sop_uid
        ds.Modality = "RTPLAN"
        ds.RTPlanLabel = "IMPT_PencilBeam"
        ds.RTPlanDate = datetime.date.today().strftime("%Y%m%d")
        ds.RTPlanTime = datetime.datetime.now().strftime("%H%M%S.%f")
        ds.RTPlanGeometry = "PATIENT"
        ds.ApprovalStatus = "UNAPPROVED"
        ds.InstanceCreationDate = ds.RTPlanDate
        ds.InstanceCreationTime = ds.RTPlanTime
        ds.Manufacturer = "RaySearch Laboratories"
        ds.PlanIntent = "CURATIVE"
        ds.TreatmentProtocols = ["PencilBeamScanning"]
        ds.PatientIdentityRemoved = "YES"
        ds.DeidentificationMethod = "RayStation Custom Export"
        ds.NumberOfBeams = len(self.beam_list)
        ds.NumberOfBrachyApplicationSetups = 0

        rss = _dcm.Dataset()
        rss.ReferencedSOPClassUID = _dcm.RT_STRUCT_STORAGE
        rss.ReferencedSOPInstanceUID = _dcm.generate_uid()
        ds.ReferencedStructureSetSequence = [rss]

        def dose_reference(number, roi, description):
            dr = _dcm.Dataset()
            dr.DoseReferenceNumber = number
            dr.ReferencedROINumber = roi
            dr.DoseReferenceStructureType = "VOLUME"
            dr.DoseReferenceDescription = description
            dr.DoseReferenceType = "TARGET"
            dr.TargetPrescriptionDose = 60
            dr.TargetUnderdoseVolumeFraction = 2
            dr.add_new((0x4001, 0x0010), "LO", "RAYSEARCHLABS 2.0")
            dr.add_new((0x4001, 0x1011), "UN", "{0}".format(60).encode())
            return dr

        if rt_struct_file:
            rt_ds = _dcm.read(rt_struct_file)
            rss.ReferencedSOPInstanceUID = rt_ds.SOPInstanceUID
            ptv_roi_num = None
            for roi in rt_ds.StructureSetROISequence:
                if roi.ROIName == "PTV":
                    ptv_roi_num = roi.ROINumber
                    break
            if ptv_roi_num is None:
                raise ValueError("No structure labelled 'PTV' found in the provided RT structure set.")
            ds.DoseReferenceSequence = [dose_reference(1, ptv_roi_num, "PTV Prescription")]
        else:
            ds.DoseReferenceSequence = [dose_reference(i + 1, i + 1, f"Dose Reference {i + 1}")
                                        for i in range(2)]

        fg = _dcm.Dataset()
        fg.FractionGroupNumber = 1
        fg.NumberOfFractionsPlanned = self.n_fractions
        fg.NumberOfBeams = len(self.beam_list)
        fg.NumberOfBrachyApplicationSetups = 0
        referenced_beams = []
        ps = _dcm.Dataset()
        ps.PatientPosition = "HFS"
        ps.PatientSetupNumber = 1
        ds.PatientSetupSequence = [ps]

        ion_beams = []
        for i, beam in enumerate(self.beam_list):
            try:
                model_index = list(self.dicom_rangeshifter_label.astype(str)).index(beam.dicom_rangeshifter_label)
            except ValueError:
                raise ValueError(f"Beam model not found for rangeshifter ID {beam.dicom_rangeshifter_label}")
            beam_model = self.beam_models[model_index]
            ion_beams.append(_ion_beam(i + 1, beam, beam_model, referenced_beams))
        fg.ReferencedBeamSequence = referenced_beams
        ds.FractionGroupSequence = [fg]
        ds.IonBeamSequence = ion_beams
        _dcm.write(ds, output_file)

    def readPlanDicom(self, plan_path):
        """Fields from an RT Ion Plan (plan_impt.py:605-643): range shifter
        label ('0' without one), gantry / couch / isocentre of the first
        control point, and the spots of every even-numbered control point,
        energies matched to the first model's labels."""
        ds = _dcm.read(plan_path, force=True)
        n_beams = len(ds.IonBeamSequence)
        self.n_fractions = float(ds.FractionGroupSequence[0].NumberOfFractionsPlanned)
        self.beam_list = []
        self.n_beams = 0
        for i in range(n_beams):
            ibs = ds.IonBeamSequence[i]
            beam = IMPTBeam()
            if ibs.NumberOfRangeShifters and ibs.RangeShifterSequence[0].RangeShifterID is not None:
                beam.dicom_rangeshifter_label = ibs.RangeShifterSequence[0].RangeShifterID
            else:
                beam.dicom_rangeshifter_label = "0"
            cp0 = ibs.IonControlPointSequence[0]
            beam.gantry_angle = float(cp0.GantryAngle)
            beam.couch_angle = float(cp0.PatientSupportAngle)
            beam.iso = np.array(cp0.IsocenterPosition, dtype=np.single)
            for j, cp in enumerate(ibs.IonControlPointSequence):
                if j % 2 == 0:
                    energy_id = self.beam_models[0].energyIDFromLabel(float(cp.NominalBeamEnergy))
                    beam.addSpotData(cp, energy_id)
            self.addBeam(beam)


def _ion_control_point(index, beam, beam_model, energy_id, spots, cumulative, weights):
    """One ion control point of an energy layer (plan_impt.py:506-581)."""
    cp = _dcm.Dataset()
    cp.ControlPointIndex = index
    cp.NominalBeamEnergyUnit = "MEV"
    cp.NominalBeamEnergy = float(beam_model.energy_labels[energy_id])
    cp.GantryAngle = float(beam.gantry_angle)
    cp.GantryRotationDirection = "NONE"
    cp.BeamLimitingDeviceAngle = 0.0
    cp.BeamLimitingDeviceRotationDirection = "NONE"
    cp.PatientSupportAngle = float(beam.couch_angle)
    cp.PatientSupportRotationDirection = "NONE"
    cp.TableTopEccentricAngle = float(beam.couch_angle)
    cp.TableTopEccentricRotationDirection = "NONE"
    cp.TableTopPitchAngle = 0.0
    cp.TableTopPitchRotationDirection = "NONE"
    cp.TableTopRollAngle = 0.0
    cp.TableTopRollRotationDirection = "NONE"
    cp.GantryPitchAngle = 0.0
    cp.GantryPitchRotationDirection = "NONE"
    cp.TableTopVerticalPosition = ""
    cp.TableTopLongitudinalPosition = ""
    cp.TableTopLateralPosition = ""
    cp.IsocenterPosition = [float(c) for c in beam.iso]
    cp.SnoutPosition = 391.0
    cp.MetersetRate = 480.0
    cp.CumulativeMetersetWeight = float(cumulative)
    cp.ScanSpotTuneID = "Standard"
    cp.NumberOfScanSpotPositions = len(spots)
    cp.ScanSpotPositionMap = spots[:, 0:2].flatten().tolist()
    cp.ScanSpotMetersetWeights = list(weights)
    width = beam_model.reference_sigmas[energy_id] * 2.0 * np.sqrt(2.0 * np.log(2.0)) * 10.0
    cp.ScanningSpotSize = [width, width]
    cp.NumberOfPaintings = 1
    if beam.dicom_rangeshifter_label:
        cp.RangeShifterSetting = "ON"
    return cp


def _ion_beam(number, beam, beam_model, referenced_beams):
    """One Ion Beam Sequence item and its fraction-group reference
    (plan_impt.py:426-601)."""
    b = _dcm.Dataset()
    b.BeamNumber = number
    b.BeamName = f"PBS_Beam{number}"
    b.BeamType = "STATIC"
    b.RadiationType = "PROTON"
    b.ScanMode = "MODULATED_SPEC"
    b.ModulatedScanModeType = "STATIONARY"
    b.TreatmentMachineName = "ProBeat_noMRF_"
    b.Manufacturer = "Generic"
    b.PrimaryDosimeterUnit = "MU"
    b.TreatmentDeliveryType = "TREATMENT"
    b.NumberOfWedges = 0
    b.NumberOfCompensators = 0
    b.NumberOfBoli = 0
    b.NumberOfBlocks = 0
    b.PatientSupportType = "TABLE"
    b.PatientSupportID = "TABLE"
    b.VirtualSourceAxisDistances = [float(beam_model.VSADX), float(beam_model.VSADY)]
    snout = _dcm.Dataset()
    snout.SnoutID = "0"
    b.SnoutSequence = [snout]
    b.NumberOfRangeShifters = 0
    if beam.dicom_rangeshifter_label:
        b.NumberOfRangeShifters = 1
        rs = _dcm.Dataset()
        rs.RangeShifterNumber = 1
        rs.RangeShifterID = beam.dicom_rangeshifter_label
        rs.RangeShifterType = "BINARY"
        rs.AccessoryCode = beam.dicom_rangeshifter_label
        b.RangeShifterSequence = [rs]
    b.NumberOfLateralSpreadingDevices = 1
    lsd = _dcm.Dataset()
    lsd.LateralSpreadingDeviceNumber = 1
    lsd.LateralSpreadingDeviceID = "0"
    lsd.LateralSpreadingDeviceType = "SCATTERER"
    b.LateralSpreadingDeviceSequence = [lsd]
    b.NumberOfRangeModulators = 1
    rm = _dcm.Dataset()
    rm.RangeModulatorNumber = 0
    rm.RangeModulatorID = "0"
    rm.RangeModulatorType = "FIXED"
    rm.RangeModulatorDescription = ""
    b.RangeModulatorSequence = [rm]

    total_beam_mu = np.sum(beam.spot_list[:, 2])
    rb = _dcm.Dataset()
    rb.ReferencedBeamNumber = number
    rb.BeamMeterset = float(total_beam_mu)
    referenced_beams.append(rb)
    b.FinalCumulativeMetersetWeight = float(total_beam_mu)
    b.add_new((0x4001, 0x0010), "LO", "RAYSEARCHLABS 2.0")
    b.add_new((0x4001, 0x1002), "UN", b"Constant 1.1")
    b.add_new((0x4001, 0x1012), "UN", b"ProBeat_noMRF_")
    b.add_new((0x4001, 0x1033), "UN", b"RBE_NOT_INCLUDED")
    b.add_new((0x4001, 0x1003), "UN", b"20181030034922.000000 ")

    lsds = _dcm.Dataset()
    lsds.LateralSpreadingDeviceSetting = "IN"
    lsds.IsocenterToLateralSpreadingDeviceDistance = 2275.0
    lsds.LateralSpreadingDeviceWaterEquivalentThickness = 0.0
    lsds.ReferencedLateralSpreadingDeviceNumber = 1
    rms = _dcm.Dataset()
    rms.ReferencedRangeModulatorNumber = 0
    rdrs = []
    for j in range(2):
        rdr = _dcm.Dataset()
        rdr.ReferencedDoseReferenceNumber = j + 1
        rdr.CumulativeDoseReferenceCoefficient = None
        rdrs.append(rdr)

    cps = []
    cumulative_mu = 0.0
    for energy_id in np.unique(beam.spot_list[:, 3]).astype(int):
        spots = beam.spot_list[beam.spot_list[:, 3] == energy_id]
        layer_mu = np.sum(spots[:, 2])
        for cumulative, weights in ((cumulative_mu, spots[:, 2].tolist()),
                                    (cumulative_mu + layer_mu, [0.0] * len(spots))):
            cp = _ion_control_point(len(cps), beam, beam_model, energy_id, spots, cumulative, weights)
            cp.LateralSpreadingDeviceSettingsSequence = [lsds]
            cp.RangeModulatorSettingsSequence = [rms]
            cp.ReferencedDoseReferenceSequence = rdrs
            cps.append(cp)
        cumulative_mu += layer_mu
    b.IonControlPointSequence = cps
    b.NumberOfControlPoints = len(cps)
    return b


def _proton_field(model, vol, beam):
    """The attribute reads of proton_raytrace_cuda and then proton_spot_cuda,
    in their order (dosemodule.cu:104-160, 243-300), so that a malformed
    input raises the error the reference's loop would raise first."""
    vsadx = dose_kernels._getfloat(model, "VSADX")
    vsady = dose_kernels._getfloat(model, "VSADY")
    dose_kernels._getarray(vol, "voxel_data", 3)
    spacing = dose_kernels._getarray(vol, "spacing", 1)
    origin = dose_kernels._getarray(vol, "origin", 1)
    iso = dose_kernels._getarray(beam, "iso", 1)
    ga = dose_kernels._getfloat(beam, "gantry_angle")
    ta = dose_kernels._getfloat(beam, "couch_angle")
    for a in ("divergence_params", "lut_depths", "lut_sigmas", "lut_idds"):
        dose_kernels._getarray(model, a, 2)
    spots = dose_kernels._getarray(beam, "spot_list", 2)
    if spots.shape[0] and spots.shape[1] < 4:
        raise ValueError("'spot_list' rows must hold x, y, MU and energy id")
    return dict(vsadx=vsadx, vsady=vsady, s=dose_kernels._spacing0(spacing),
                origin=dose_kernels._vec3(origin, "origin"), iso=dose_kernels._vec3(iso, "iso"),
                ga=ga, ta=ta, rows=np.ascontiguousarray(spots[:, :4], dtype=np.float32))


class IMPTDoseGrid(DoseGrid):

    def __init__(self):
        super().__init__()
        self.RLSP = []

    def RLSPFromHU(self, machine_name):
        return _bm.rlsp_from_hu(self.HU, machine_name)

    def _rlsp_object(self):
        return _volume_object(self.RLSP, self.origin, self.spacing)

    def computeIMPTPlan(self, plan, gpu_id=0):
        """Sum of the fields' pencil-beam doses times the fraction count
        (plan_impt.py:124-160), with the RLSP uploaded once and every WET
        and dose kept on the device until the end."""
        from . import engine
        self.beam_doses = []
        self.dose = np.zeros(self.size, dtype=np.single)
        table = _bm.hu_table("protons", plan.machine_name)
        self._check_isotropic()
        vol = _checks_volume(self.HU, self.origin, self.spacing)
        res = _Resident(self.HU, gpu_id, table)
        beam_ts = []
        for beam in plan.beam_list:
            model = plan.beam_models[_model_index(plan.dicom_rangeshifter_label,
                                                  beam.dicom_rangeshifter_label,
                                                  "rangeshifter ID")]
            f = _proton_field(model, vol, beam)
            rlsp_t = res.up().vol
            wet_t = engine.proton_wet(rlsp_t, f["s"], f["origin"], f["iso"], f["ga"], f["ta"],
                                      f["vsadx"], f["vsady"], stream=res.stream)
            pp = engine.ProtonPlan(res.dev.index, rlsp_t.shape, f["s"], f["origin"], f["iso"],
                                   f["ga"], f["ta"], model, f["rows"])
            try:
                beam_t = pp.dose(wet_t, stream=res.stream)
            finally:
                # the field's spot and layer buffers are freed by close()
                res.stream.synchronize()
                pp.close()
            beam_ts.append(beam_t)
            res.sum.add_(beam_t)
        self.RLSP = res.host() if res.torch is not None else self.RLSPFromHU(plan.machine_name)
        self._finish(res, beam_ts, plan.n_fractions, scale_beams=True)

    def computeIMPTWET(self, plan, gpu_id=0):
        """One smoothed-WET volume per field (plan_impt.py:162-190); the RLSP
        is converted on the device at the first field."""
        table = _bm.hu_table("protons", plan.machine_name)
        res = _Resident(self.HU, gpu_id, table)
        vol = None
        beam_wets = []
        for beam in plan.beam_list:
            model = plan.beam_models[_model_index(plan.dicom_rangeshifter_label,
                                                  beam.dicom_rangeshifter_label,
                                                  "rangeshifter ID")]
            if vol is None:
                self.RLSP = res.host()
                vol = self._rlsp_object()
            beam_wet = dose_kernels.proton_raytrace_cuda(model, vol, beam, gpu_id)
            beam_wets.append(_volume_object(beam_wet, self.origin, self.spacing))
        if vol is None:
            self.RLSP = self.RLSPFromHU(plan.machine_name)
        return beam_wets

    def computeIMPTPlanFromWET(self, plan, wet_object, beam, gpu_id=0):
        """One field's dose from a WET volume computed earlier
        (plan_impt.py:192-211); uses the RLSP of the last computeIMPTWET."""
        self._check_isotropic()
        vol = self._rlsp_object()
        model = plan.beam_models[_model_index(plan.dicom_rangeshifter_label,
                                              beam.dicom_rangeshifter_label, "rangeshifter ID")]
        return dose_kernels.proton_spot_cuda(model, vol, wet_object, beam, gpu_id)

    def writeWETNRRD(self, plan, wet_path, gpu_id=0):
        """One WET file per field, ``x_beamNN.nrrd`` (plan_impt.py:213-242)."""
        if not wet_path.endswith(".nrrd"):
            raise Exception("WET path must have .nrrd extension")
        for i, wet in enumerate(self.computeIMPTWET(plan, gpu_id)):
            _nrrd.write(_nrrd.beam_path(wet_path, i), wet.voxel_data, self.origin, self.spacing)


# ---------------------------------------------------------------------------
# photons (plan_imrt.py)
# ---------------------------------------------------------------------------
class IMRTPhotonEnergy(_bm.PhotonBeamModel):
    """One photon energy of a machine (plan_imrt.py:13-92).  Constructed
    empty, as in the reference; IMRTPlan._load_beam_model_parameters fills it
    from ``lookuptables/photons/<machine>/<folder>``."""

    _FIELDS = ("output_factor_equivalent_squares", "output_factor_values", "mu_calibration",
               "primary_source_distance", "scatter_source_distance", "mlc_distance",
               "scatter_source_weight", "electron_attenuation", "primary_source_size",
               "scatter_source_size", "profile_radius", "profile_intensities",
               "profile_softening", "spectrum_attenuation_coefficients",
               "spectrum_primary_weights", "spectrum_scatter_weights", "electron_source_weight",
               "has_xjaws", "has_yjaws", "electron_fitted_dmax", "jaw_transmission",
               "mlc_transmission", "mlc_index", "mlc_widths", "mlc_offsets", "n_mlc_pairs",
               "kernel")

    def __init__(self, dicom_energy_label):
        self.dicom_energy_label = dicom_energy_label
        for k in self._FIELDS:
            setattr(self, k, None)

    def validate_parameters(self):
        for param in self._FIELDS[:22]:
            if getattr(self, param, None) is None:
                raise Exception(f"{param} not set in beam model")


class IMRTControlPoint(_bm.PhotonControlPoint):

    def __init__(self, iso, mu, mlc, ga, ca, ta, xjaws, yjaws):
        self.iso = iso
        self.mu = mu
        self.mlc = mlc
        self.ga = ga
        self.ca = ca
        self.ta = ta
        self.xjaws = xjaws
        self.yjaws = yjaws


class IMRTBeam(Beam):

    def __init__(self):
        super().__init__()
        self.cp_list = []
        self.n_cps = 0
        self.dicom_energy_label = None

    def addControlPoint(self, cp):
        self.cp_list.append(cp)
        self.n_cps += 1


class IMRTPlan(Plan):

    def __init__(self, machine_name="VarianTrueBeamHF"):
        super().__init__()
        self.machine_name = machine_name
        t = _bm._read_table(os.path.join(_bm.TABLES, "photons", machine_name, "energy_labels.csv"))
        self.dicom_energy_label = np.array(t["dicom_energy_label"])
        self.folder_energy_label = np.array(t["folder_energy_label"])
        self.beam_models = []
        for d, f in zip(self.dicom_energy_label, self.folder_energy_label):
            beam_model = IMRTPhotonEnergy(d)
            self._load_beam_model_parameters(beam_model, machine_name, f)
            self.beam_models.append(beam_model)

    def _load_beam_model_parameters(self, beam_model, machine_name, folder_energy_label):
        """Machine geometry, MLC geometry, kernel and beam parameters of one
        energy (plan_imrt.py:192-271), then the completeness check."""
        _bm.PhotonBeamModel.__init__(beam_model, machine_name, folder_energy_label,
                                     beam_model.dicom_energy_label)
        beam_model.validate_parameters()

    def readPlanDicom(self, plan_path):
        """Treatment beams of an RT Plan (plan_imrt.py:273-351).  Each control
        point after the first becomes a control point with the aperture and
        angles in force before it and MU = (its cumulative meterset weight -
        the previous one) x the first referenced beam's meterset."""
        ds = _dcm.read(plan_path)
        group = ds.FractionGroupSequence[0]
        meterset = group.ReferencedBeamSequence[0].BeamMeterset
        self.n_fractions = float(group.NumberOfFractionsPlanned)
        self.beam_list = []
        self.n_beams = 0
        for item in ds.BeamSequence:
            if item.TreatmentDeliveryType == "TREATMENT":
                self.addBeam(_beam_from_control_points(item.ControlPointSequence, self.beam_models[0], meterset))

    def addSquareField(self, dicom_energy_label='6', dimx=10, dimy=10, mu=100, gantry_angle=0.0,
                       collimator_angle=0.0, table_angle=0.0):
        """One beam with one control point of a dimx x dimy cm field at the
        origin (plan_imrt.py:353-385); MLC geometry of the first model."""
        imrt_beam = IMRTBeam()
        imrt_beam.dicom_energy_label = dicom_energy_label
        iso = np.array([0.0, 0.0, 0.0], dtype=np.single)
        mlc = _bm.square_field_mlc(self.beam_models[0], dimx, dimy)
        jawx = np.array([-dimx * 10.0 / 2.0, dimx * 10.0 / 2.0], dtype=np.single)
        jawy = np.array([-dimy * 10.0 / 2.0, dimy * 10.0 / 2.0], dtype=np.single)
        cp = IMRTControlPoint(iso, mu, mlc, gantry_angle, collimator_angle, table_angle, jawx, jawy)
        imrt_beam.addControlPoint(cp)
        self.addBeam(imrt_beam)


def _limiting_devices(cp):
    """{device type: positions} of a control point's Beam Limiting Device
    Position Sequence, read where the reference reads them
    (plan_imrt.py:298-306): an X or a Y device takes the positions of the
    sequence's item 1 and an MLCX device those of item 2, whatever those
    items' own types."""
    seq = cp.BeamLimitingDevicePositionSequence
    item_of = {"X": 1, "Y": 1, "MLCX": 2}
    kinds = [ps.RTBeamLimitingDeviceType for ps in seq]
    return {k: np.array(seq[item_of[k]].LeafJawPositions, dtype=np.single) for k in item_of if k in kinds}


class _Segment:
    """Machine state between two control points: angles, beam limiting
    devices and the cumulative meterset weight reached."""

    def __init__(self, cp):
        self.ga = cp.GantryAngle
        self.ca = cp.BeamLimitingDeviceAngle
        self.ta = cp.PatientSupportAngle
        self.weight = 0.0
        self._devices(cp)

    def _devices(self, cp):
        self.dev = _limiting_devices(cp)
        self.used = False

    def aperture(self, model):
        """Rows (x1, x2, y offset, width) of the leaves in force, or of the
        jaws when there is no MLC: the pairs span the X jaws (closed without
        them) and close outside the Y jaws."""
        n = model.n_mlc_pairs
        if self.used:
            # the reference keeps the rows it built in place of the leaf
            # positions, so a second segment without new device positions
            # fails reshaping them, as it does there
            np.reshape(np.zeros(4 * n, np.single), (2, n))
        self.used = True
        if "MLCX" in self.dev:
            banks = np.reshape(self.dev["MLCX"], (2, n))
            return aperture_rows(model, banks[0], banks[1], np.ones(n, bool))
        x, y = self.dev.get("X"), self.dev.get("Y")
        off = np.asarray(model.mlc_offsets)
        inside = np.ones(n, bool) if y is None else ~((off < y[0]) | (off > y[1]))
        if x is None:
            return aperture_rows(model, 0.0, 0.0, np.zeros(n, bool))
        return aperture_rows(model, x[0], x[1], inside)

    def advance(self, cp):
        if "GantryAngle" in cp:
            self.ga = cp.GantryAngle
        if "BeamLimitingDevicePositionSequence" in cp:
            self._devices(cp)
        self.weight = cp.CumulativeMetersetWeight


def _beam_from_control_points(sequence, model, meterset):
    """An IMRTBeam from an RT Plan control point sequence: control point 0
    sets the energy, isocentre, angles and devices; every later control point
    closes a segment (MU = weight step x meterset) with the state in force
    before it, then applies its own gantry angle and devices."""
    beam = IMRTBeam()
    seg = None
    for cp in sequence:
        if cp.ControlPointIndex == 0:
            beam.dicom_energy_label = cp.NominalBeamEnergy
            beam.iso = np.array([cp.IsocenterPosition[a] for a in range(3)], dtype=np.single)
            seg = _Segment(cp)
            continue
        mu = (cp.CumulativeMetersetWeight - seg.weight) * meterset
        beam.addControlPoint(IMRTControlPoint(beam.iso, mu, seg.aperture(model), seg.ga, seg.ca, seg.ta,
                                              seg.dev.get("X"), seg.dev.get("Y")))
        seg.advance(cp)
    return beam


class IMRTDoseGrid(DoseGrid):

    def __init__(self):
        super().__init__()
        self.Density = []

    def DensityFromHU(self, machine_name):
        return _bm.density_from_hu(self.HU, machine_name)

    def computeIMRTPlan(self, plan, gpu_id=0):
        """Sum over beams of Σ_cp outputFactor(cp) · CCCS dose(cp), times the
        fraction count (plan_imrt.py:126-158).  Every control point is read
        and checked first (the reference's reads, in its order); then the
        density is converted on the device and the beams run as cases of the
        batching scheduler (two streams when there are several beams), each
        accumulating its control points' doses times their float32 output
        factors in order, so the per-beam and plan sums equal the reference
        loop's bit for bit.  Control points of a beam with the same source
        position share one ray trace."""
        from . import engine
        from .scheduler import Scheduler
        self.beam_doses = []
        self.dose = np.zeros(self.size, dtype=np.single)
        table = _bm.hu_table("photons", plan.machine_name)
        self._check_isotropic()
        vol = _checks_volume(self.HU, self.origin, self.spacing)
        res = _Resident(self.HU, gpu_id, table)
        beams = []
        for beam in plan.beam_list:
            model = plan.beam_models[_model_index(plan.dicom_energy_label,
                                                  beam.dicom_energy_label, "beam energy")]
            params = None
            cps = []
            for cp in beam.cp_list:
                output_factor = model.outputFactor(cp)
                # photon_dose_cuda's reads in its order (dosemodule.cu:386-460):
                # model, volume, control point
                if params is None:
                    params = engine.PhotonParams(model)
                dose_kernels._getarray(vol, "voxel_data", 3)
                spacing = dose_kernels._getarray(vol, "spacing", 1)
                origin = dose_kernels._getarray(vol, "origin", 1)
                iso = dose_kernels._getarray(cp, "iso", 1)
                mlc = dose_kernels._getarray(cp, "mlc", 2)
                mu = dose_kernels._getfloat(cp, "mu")
                ga = dose_kernels._getfloat(cp, "ga")
                ca = dose_kernels._getfloat(cp, "ca")
                ta = dose_kernels._getfloat(cp, "ta")
                if mlc.shape[0] and mlc.shape[1] < 4:
                    raise ValueError("'mlc' rows must hold x1, x2, y_offset and y_width")
                # float32 product then float32 sum, as numpy < 2 evaluates
                # beam_dose += cp_dose * output_factor
                cps.append((dose_kernels._spacing0(spacing), dose_kernels._vec3(origin, "origin"),
                            dose_kernels._vec3(iso, "iso"), mu, ga, ca, ta, mlc,
                            float(np.float32(output_factor))))
            beams.append((params, cps))
        beam_ts = [res.zeros() for _ in beams]
        busy = [i for i, (_, cps) in enumerate(beams) if cps]
        if busy:
            sched = Scheduler(res.dev.index, streams=min(2, len(busy)))
            try:
                for i in busy:
                    params, cps = beams[i]
                    case = sched.add_case(res.vol, cps[0][0], cps[0][1], dose=beam_ts[i])
                    for sp, org, iso, mu, ga, ca, ta, mlc, of in cps:
                        sched.add_photon(case, iso, mu, ga, ca, ta, params, mlc, scale=of)
                sched.run(graph=False, stream=res.stream)
            finally:
                sched.close()
        for beam_t in beam_ts:
            res.sum.add_(beam_t)
        self.Density = res.host() if res.torch is not None else self.DensityFromHU(plan.machine_name)
        self._finish(res, beam_ts, plan.n_fractions, scale_beams=False)
