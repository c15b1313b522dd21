"""Beam-model loaders for the unchanged lookup tables (dosecuda_amd/lookuptables,
the reference's DoseCUDA/lookuptables data), plus the caller-side helpers on
either side of the engine: HU conversion, square-field MLC, output factor.

These restate the parsing rules of the reference's plan code so that the
engine receives exactly the arrays the reference's callers would pass:
  IMPTBeamModel.__init__      DoseCUDA/plan_impt.py:17-94
  RLSPFromHU                  DoseCUDA/plan_impt.py:112-122
  DensityFromHU               DoseCUDA/plan_imrt.py:114-124
  _load_beam_model_parameters DoseCUDA/plan_imrt.py:192-271
  addSquareField (MLC rule)   DoseCUDA/plan_imrt.py:353-382
  outputFactor                DoseCUDA/plan_imrt.py:61-92
pandas is not needed: the CSVs are read with the csv module and numpy.
"""
from __future__ import annotations

import csv
import os

import numpy as np

TABLES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lookuptables")


def _read_table(path):
    with open(path, newline="") as f:
        rows = [r for r in csv.reader(f) if r]
    head = [h.strip() for h in rows[0]]
    cols = {h: [] for h in head}
    for r in rows[1:]:
        for h, v in zip(head, r):
            cols[h].append(v.strip())
    return cols


class ProtonBeamModel:
    """Attributes read by dose_kernels.proton_*: VSADX, VSADY,
    divergence_params [n_E, 5], lut_depths / lut_sigmas / lut_idds [n_E, 400]."""

    def __init__(self, machine="HitachiProbeatJHU", rangeshifter_folder="RS0",
                 dicom_rangeshifter_label="0"):
        root = os.path.join(TABLES, "protons", machine)
        self.dicom_rangeshifter_label = dicom_rangeshifter_label
        self.VSADX = self.VSADY = None
        for line in open(os.path.join(root, "machine_geometry.csv")):
            if line.startswith("VSADX"):
                self.VSADX = float(line.split(",")[1])
            if line.startswith("VSADY"):
                self.VSADY = float(line.split(",")[1])
        if self.VSADX is None or self.VSADY is None:
            raise Exception("VSADX/VSADY not found in machine_geometry.csv")
        et = _read_table(os.path.join(root, rangeshifter_folder, "energies.csv"))
        self.energy_labels = np.array([float(v) for v in et["energy_label"]])
        ids = [int(v) for v in et["index"]]
        self.reference_sigmas = np.array([float(v) for v in et["sigma"]])
        dvp, dep, sig, idd = [], [], [], []
        for eid in ids:
            path = os.path.join(root, rangeshifter_folder, "energy_%03d.csv" % eid)
            with open(path) as f:
                f.readline()                                   # header
                dvp.append([float(p) for p in f.readline().split(",")])
                f.readline()                                   # blank
                f.readline()                                   # header
                d, s, i = [], [], []
                for k, line in enumerate(f):
                    if k > 399:                                # plan_impt.py:73-75
                        break
                    p = line.split(",")
                    d.append(float(p[0])); s.append(float(p[1])); i.append(float(p[2]))
            dep.append(d); sig.append(s); idd.append(i)
        self.divergence_params = np.array(dvp, dtype=np.single)
        self.lut_depths = np.array(dep, dtype=np.single)
        self.lut_sigmas = np.array(sig, dtype=np.single)
        self.lut_idds = np.array(idd, dtype=np.single)

    def energyIDFromLabel(self, energy_label):
        hits = np.nonzero(self.energy_labels == energy_label)[0]
        if hits.size != 1:
            raise Exception("Energy ID not found for %.2f" % energy_label)
        return int(hits[0])


def hu_table(kind="protons", machine="HitachiProbeatJHU"):
    name = "HU_RLSP.csv" if kind == "protons" else "HU_Density.csv"
    t = _read_table(os.path.join(TABLES, kind, machine, name))
    val = "RLSP" if kind == "protons" else "Density"
    return (np.array([float(v) for v in t["HU"]]), np.array([float(v) for v in t[val]]))


def rlsp_from_hu(hu, machine="HitachiProbeatJHU"):
    """plan_impt.py:112-122: np.interp with end clamping, then float32."""
    h, r = hu_table("protons", machine)
    return np.array(np.interp(hu, h, r), dtype=np.single)


def density_from_hu(hu, machine="VarianTrueBeamHF"):
    """plan_imrt.py:114-124."""
    h, d = hu_table("photons", machine)
    return np.array(np.interp(hu, h, d), dtype=np.single)


class PhotonBeamModel:
    """Attributes read by dose_kernels.photon_dose_cuda (dosemodule.cu:400-437)."""

    def __init__(self, machine="VarianTrueBeamHF", energy_folder="6MV", dicom_energy_label="6"):
        root = os.path.join(TABLES, "photons", machine)
        self.dicom_energy_label = dicom_energy_label
        g = _read_table(os.path.join(root, "mlc_geometry.csv"))
        self.mlc_index = np.array([int(v) for v in g["mlc_pair_index"]])
        self.mlc_widths = np.array([float(v) for v in g["width"]])
        self.mlc_offsets = np.array([float(v) for v in g["center_offset"]])
        self.n_mlc_pairs = len(self.mlc_index)
        kt = _read_table(os.path.join(root, energy_folder, "kernel.csv"))
        cols = ["theta", "Am", "am", "Bm", "bm", "ray_length"]
        k = np.array([[float(v) for v in kt[c]] for c in cols], dtype=np.single).T
        # pandas' DataFrame.to_numpy() of a float frame is column-major; the
        # reference reads the buffer column by column (IMRTClasses.cu:290-296)
        self.kernel = np.asfortranarray(k)
        for line in open(os.path.join(root, "machine_geometry.csv")):
            key, _, val = line.partition(",")
            if key == "primary_source_distance":
                self.primary_source_distance = float(val)
            elif key == "scatter_source_distance":
                self.scatter_source_distance = float(val)
            elif key == "mlc_distance":
                self.mlc_distance = float(val)
            elif key == "has_xjaws":
                self.has_xjaws = bool(val)        # plan_imrt.py:230-233 (bool of a string)
            elif key == "has_yjaws":
                self.has_yjaws = bool(val)
        arr_keys = ("output_factor_equivalent_squares", "output_factor_values", "profile_radius",
                    "profile_intensities", "profile_softening",
                    "spectrum_attenuation_coefficients", "spectrum_primary_weights",
                    "spectrum_scatter_weights")
        for line in open(os.path.join(root, energy_folder, "beam_parameters.csv")):
            parts = line.split(",")
            key = parts[0]
            if key in arr_keys:
                setattr(self, key, np.array([float(p) for p in parts[1:]], dtype=np.single))
            elif key in ("mu_calibration", "scatter_source_weight", "electron_attenuation",
                         "primary_source_size", "scatter_source_size", "electron_source_weight",
                         "electron_fitted_dmax", "jaw_transmission", "mlc_transmission"):
                setattr(self, key, float(parts[1]))

    def outputFactor(self, cp):
        """Equivalent-square output factor of a control point's aperture
        (plan_imrt.py:61-92): 4 * area / perimeter, where the area sums every
        pair's opening times its width and the perimeter spans the y extent
        and the widest opening of the pairs open by more than 3 mm, looked up
        in the model's output-factor table.

        The arithmetic is the reference's under numpy < 2 (its pinned
        environment), written out: a pair's opening and area are float32
        (float32 scalar operations), the running area, the pair edges
        (offset -/+ width / 2.0) and the perimeter are float64, and the areas
        are added in pair order."""
        rows = np.asarray(cp.mlc)
        x1, x2 = rows[:, 0].astype(np.single), rows[:, 1].astype(np.single)
        off, wid = rows[:, 2].astype(np.single), rows[:, 3].astype(np.single)
        opening = x2 - x1                                   # float32
        area = 0.0
        for a in (opening * wid).tolist():                  # float32 products, float64 sum
            area += a
        big = opening > 3.0
        lo, hi, widest = 10000.0, -10000.0, 0.0
        if big.any():
            half = wid[big].astype(np.float64) / 2.0
            lo = min(lo, float((off[big] - half).min()))
            hi = max(hi, float((off[big] + half).max()))
            widest = float(opening[big].max())
        perimeter = 2.0 * (hi - lo) + 2.0 * widest
        return np.interp(4.0 * area / perimeter, self.output_factor_equivalent_squares,
                         self.output_factor_values)


def aperture_rows(model, x_lo, x_hi, open_pairs):
    """[n_pairs, 4] float32 rows (x1, x2, y offset, width), column-major as
    the reference's transposed stack: pairs in open_pairs span [x_lo, x_hi]
    (scalars or per-pair arrays, in mm), the others are closed at 0."""
    n = model.n_mlc_pairs
    rows = np.zeros((n, 4), dtype=np.single, order="F")
    rows[:, 0] = np.where(open_pairs, x_lo, 0.0)
    rows[:, 1] = np.where(open_pairs, x_hi, 0.0)
    rows[:, 2] = model.mlc_offsets
    rows[:, 3] = model.mlc_widths
    return rows


def square_field_mlc(model: PhotonBeamModel, dimx_cm=10.0, dimy_cm=10.0):
    """Aperture of addSquareField (plan_imrt.py:360-374): a dimx x dimy cm
    field centred on the axis; a pair opens from -X/2 to +X/2 (X in mm)
    when its centre offset lies within +-Y/2, otherwise it stays closed."""
    half_x, half_y = dimx_cm * 10.0 / 2.0, dimy_cm * 10.0 / 2.0
    off = np.asarray(model.mlc_offsets)
    return aperture_rows(model, -half_x, half_x, (off >= -half_y) & (off <= half_y))


class VolumeObject:
    """plan.py:6-11: voxel_data (nz, ny, nx) float32, origin, spacing."""

    def __init__(self, voxel_data=None, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0)):
        self.voxel_data = voxel_data if voxel_data is not None else []
        self.origin = np.array(origin, dtype=np.single)
        self.spacing = np.array(spacing, dtype=np.single)


class ProtonField:
    """Beam attributes read by the proton entry points."""

    def __init__(self, iso, gantry_angle, couch_angle, spot_list, dicom_rangeshifter_label="0"):
        self.iso = np.array(iso, dtype=np.single)
        self.gantry_angle = gantry_angle
        self.couch_angle = couch_angle
        self.spot_list = np.array(spot_list, dtype=np.single).reshape(-1, 4)
        self.dicom_rangeshifter_label = dicom_rangeshifter_label


class PhotonControlPoint:
    """IMRTControlPoint (plan_imrt.py:95-105)."""

    def __init__(self, iso, mu, mlc, ga, ca, ta, xjaws=None, yjaws=None):
        self.iso = np.array(iso, dtype=np.single)
        self.mu = mu
        self.mlc = mlc
        self.ga, self.ca, self.ta = ga, ca, ta
        self.xjaws, self.yjaws = xjaws, yjaws
