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

The DICOM and resampling methods need pydicom / SimpleITK, which this build
does not depend on; they raise NotImplementedError naming the missing
package.  NRRD files are read and written by ``dosecuda_amd.nrrd``.
"""
from __future__ import annotations

import os
import sys

import numpy as np

from . import beam_models as _bm
from . import dose_kernels
from . import nrrd as _nrrd
from .beam_models import VolumeObject

__all__ = ["VolumeObject", "DoseGrid", "Beam", "Plan", "IMPTBeamModel", "IMPTDoseGrid",
           "IMPTBeam", "IMPTPlan", "IMRTPhotonEnergy", "IMRTControlPoint", "IMRTDoseGrid",
           "IMRTBeam", "IMRTPlan"]


def _needs(pkg: str, what: str):
    raise NotImplementedError("%s needs %s, which this build does not depend on" % (what, pkg))


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
    are raised here before it too."""

    def __init__(self, data, gpu_id):
        self.data, self.gpu_id, self.torch = data, gpu_id, None

    def up(self):
        if self.torch is None:
            self.torch, self.dev = _torch_device(self.gpu_id)
            self.vol = self.torch.from_numpy(np.ascontiguousarray(self.data)).to(self.dev)
            self.sum = self.torch.zeros_like(self.vol)
            # torch's elementwise work on these tensors runs on this stream
            # too, so the engine's launches and the sums stay in order
            self.stream = self.torch.cuda.current_stream(self.dev)
        return self

    def zeros(self):
        return self.up().torch.zeros_like(self.vol)


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
        _needs("pydicom and SimpleITK", "loadCTDCM")

    def resampleCT(self, new_spacing, new_size, new_origin):
        _needs("SimpleITK", "resampleCT")

    def resampleCTfromSpacing(self, spacing):
        _needs("SimpleITK", "resampleCTfromSpacing")

    def resampleCTfromReferenceDose(self, ref_dose_path):
        _needs("pydicom and SimpleITK", "resampleCTfromReferenceDose")

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
        _needs("pydicom", "writeDoseDCM")

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
        _needs("SimpleITK", "writeCTNIFTI")

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
        _needs("pydicom", "load_ct_series")

    def writePlanDicom(self, ct_dir, output_file, rt_struct_file=None):
        _needs("pydicom", "writePlanDicom")

    def readPlanDicom(self, plan_path):
        _needs("pydicom", "readPlanDicom")


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
        self.RLSP = self.RLSPFromHU(plan.machine_name)
        self._check_isotropic()
        vol = self._rlsp_object()
        res = _Resident(vol.voxel_data, gpu_id)
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
        self._finish(res, beam_ts, plan.n_fractions, scale_beams=True)

    def computeIMPTWET(self, plan, gpu_id=0):
        """One smoothed-WET volume per field (plan_impt.py:162-190)."""
        self.RLSP = self.RLSPFromHU(plan.machine_name)
        vol = self._rlsp_object()
        beam_wets = []
        for beam in plan.beam_list:
            model = plan.beam_models[_model_index(plan.dicom_rangeshifter_label,
                                                  beam.dicom_rangeshifter_label,
                                                  "rangeshifter ID")]
            beam_wet = dose_kernels.proton_raytrace_cuda(model, vol, beam, gpu_id)
            beam_wets.append(_volume_object(beam_wet, self.origin, self.spacing))
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
        _needs("pydicom", "readPlanDicom")

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


class IMRTDoseGrid(DoseGrid):

    def __init__(self):
        super().__init__()
        self.Density = []

    def DensityFromHU(self, machine_name):
        return _bm.density_from_hu(self.HU, machine_name)

    def computeIMRTPlan(self, plan, gpu_id=0):
        """Sum over beams of Σ_cp outputFactor(cp) · CCCS dose(cp), times the
        fraction count (plan_imrt.py:126-158).  The density is uploaded once;
        each control point's dose is scaled and added on the device."""
        from . import engine
        self.beam_doses = []
        self.dose = np.zeros(self.size, dtype=np.single)
        self.Density = self.DensityFromHU(plan.machine_name)
        self._check_isotropic()
        vol = _volume_object(self.Density, self.origin, self.spacing)
        res = _Resident(vol.voxel_data, gpu_id)
        beam_ts = []
        cp_t = None
        for beam in plan.beam_list:
            model = plan.beam_models[_model_index(plan.dicom_energy_label,
                                                  beam.dicom_energy_label, "beam energy")]
            params = None
            beam_t = None
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
                if beam_t is None:
                    beam_t = res.zeros()
                if cp_t is None:
                    cp_t = res.zeros()
                engine.photon_dose(res.vol, dose_kernels._spacing0(spacing),
                                   dose_kernels._vec3(origin, "origin"),
                                   dose_kernels._vec3(iso, "iso"), mu, ga, ca, ta, params, mlc,
                                   out=cp_t, stream=res.stream)
                # float32 product then float32 sum, as numpy < 2 evaluates
                # beam_dose += cp_dose * output_factor
                cp_t.mul_(float(np.float32(output_factor)))
                beam_t.add_(cp_t)
            beam_t = res.zeros() if beam_t is None else beam_t
            beam_ts.append(beam_t)
            res.sum.add_(beam_t)
        self._finish(res, beam_ts, plan.n_fractions, scale_beams=False)
