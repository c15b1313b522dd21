"""ctypes binding of libdosehip.so (include/dosehip.h).

The engine has no CPU fallback: if the library is missing or cannot be
loaded, every entry point raises.  torch (when importable) is imported first
so that the engine and torch share one HIP runtime in the process.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

try:  # share torch's HIP runtime when torch is present (same SONAME)
    import torch  # noqa: F401
except Exception:  # pragma: no cover - torch is optional for the C ABI
    torch = None

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdosehip.so")

DH_OK = 0
DH_E_INVALID = -1
DH_E_RUNTIME = -2
DH_E_NOMEM = -3
LUT_LENGTH = 400


class Grid(C.Structure):
    _fields_ = [("nz", C.c_int32), ("ny", C.c_int32), ("nx", C.c_int32),
                ("spacing", C.c_float), ("origin", C.c_float * 3)]


class ProtonBeam(C.Structure):
    _fields_ = [("iso", C.c_float * 3), ("gantry_angle", C.c_double),
                ("couch_angle", C.c_double), ("vsadx", C.c_double), ("vsady", C.c_double)]


class ProtonLut(C.Structure):
    _fields_ = [("n_energies", C.c_int32), ("row_stride", C.c_int32),
                ("divergence_params", C.c_void_p), ("depths", C.c_void_p),
                ("sigmas", C.c_void_p), ("idds", C.c_void_p)]


class PhotonCP(C.Structure):
    _fields_ = [("iso", C.c_float * 3), ("mu", C.c_double), ("ga", C.c_double),
                ("ca", C.c_double), ("ta", C.c_double)]


class PhotonModel(C.Structure):
    _fields_ = [("n_profile", C.c_int32), ("profile_radius", C.c_void_p),
                ("profile_intensities", C.c_void_p), ("profile_softening", C.c_void_p),
                ("n_spectral", C.c_int32), ("spectrum_attenuation_coefficients", C.c_void_p),
                ("spectrum_primary_weights", C.c_void_p), ("spectrum_scatter_weights", C.c_void_p),
                ("mu_calibration", C.c_double), ("primary_source_distance", C.c_double),
                ("scatter_source_distance", C.c_double), ("primary_source_size", C.c_double),
                ("scatter_source_size", C.c_double), ("mlc_distance", C.c_double),
                ("scatter_source_weight", C.c_double), ("electron_attenuation", C.c_double),
                ("electron_source_weight", C.c_double), ("electron_fitted_dmax", C.c_double),
                ("jaw_transmission", C.c_double), ("mlc_transmission", C.c_double),
                ("has_xjaws", C.c_int32), ("has_yjaws", C.c_int32), ("kernel", C.c_float * 36)]


_P = C.c_void_p
_I = C.c_int
# symbol -> argtypes (every symbol declared in include/dosehip.h)
SIGNATURES = {
    "dh_last_error": [],
    "dh_version": [],
    "dh_device_count": [C.POINTER(C.c_int)],
    "dh_init": [_I],
    "dh_release": [_I],
    "dh_proton_raytrace": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), _P, _P],
    "dh_proton_spot": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), C.POINTER(ProtonLut), _P,
                       C.c_int32, _P, _P],
    "dh_photon_dose": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P,
                       C.c_int32, _P, _P, _P, _P],
    "dh_proton_raytrace_dev": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), _P, _P, _P, _P],
    "dh_proton_raytrace_dev_noalloc": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), _P, _P, _P],
    "dh_proton_plan_create": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), C.POINTER(ProtonLut),
                              _P, C.c_int32, C.POINTER(C.c_void_p)],
    "dh_proton_plan_layers": [_P, C.POINTER(C.c_int32), _P, _P],
    "dh_proton_plan_dose_dev": [_P, _P, _P, _P],
    "dh_proton_plan_destroy": [_P],
    "dh_photon_dose_dev": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P,
                           C.c_int32, _P, _P, _P, _P, _P, _P],
    "dh_photon_terma_dev": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P,
                            C.c_int32, _P, C.c_float, _P, _P, _P],
    "dh_photon_terma_batch_dev": [_I, C.POINTER(Grid), C.c_int32, C.POINTER(PhotonCP), C.POINTER(PhotonModel),
                                  _P, _P, _P, C.c_float, _P, _P, _P],
    "dh_hu_convert_dev": [_I, _P, C.c_int32, C.c_int64, _P, _P, C.c_int32, _P, _P],
    "dh_stage_raytrace_dev": [_I, C.POINTER(Grid), _P, C.c_double, C.c_double, C.c_float,
                              C.c_float, _P, _P, _P],
    "dh_stage_smooth_dev": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), _P, _P, _P],
    "dh_stage_terma_dev": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P,
                           C.c_int32, _P, _P, _P, _P],
    "dh_stage_ccc_dev": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, _P,
                         _P, _P, _P],
    "dh_influence_create": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), C.POINTER(ProtonLut), _P,
                            C.c_int32, C.c_int32, C.c_int32, C.c_float, C.POINTER(C.c_void_p)],
    "dh_influence_wet": [_P, _P, C.c_float, _P, _P],
    "dh_influence_footprint": [_P, _P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)],
    "dh_influence_count": [_P, _P, C.c_int64, C.POINTER(C.c_int64), _P, _P],
    "dh_influence_count_many": [_P, C.c_int32, _P, C.c_float, _P, C.c_int64, _P, _P, _P],
    "dh_influence_fill": [_P, _P, _P, _P, _P],
    "dh_influence_info": [_P, C.POINTER(C.c_int32), _P, _P],
    "dh_influence_destroy": [_P],
    "dh_influence_stats": [_P, _P, _P],
    "dh_stage_ccc_list_dev": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P,
                              _P, _P, _P, C.c_int32, _P, _P],
    "dh_nn_conv3d_dev": [_P, C.c_int32, C.c_int32, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                         C.c_int32, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                         _P, _P, C.c_int32, C.c_int32, _P],
    "dh_nn_pool2_dev": [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P],
    "dh_sur_pack_dev": [C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, C.c_float, _P, _P, _P],
    "dh_sur_finish_dev": [C.c_int32, C.c_int32, C.c_int32, _P, _P, C.c_float, _P, _P, C.c_float,
                          _P, _P, _P],
    "dh_sur_check_work_size": [C.POINTER(Grid), C.c_int32, C.POINTER(C.c_int64)],
    "dh_sur_check_dev": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, _P,
                         _P, _P, _P, C.c_int32, C.c_float, C.c_int32, _P, C.c_int64, _P, _P],
    "dh_nn_conv3d_wgrad_dev": [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                               C.c_int32, C.c_int32, _P, _P],
    "dh_nn_prep_weights_dev": [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P],
    "dh_gamma_dev": [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                     C.c_float, C.c_int32, _P, _P],
    "dh_sur_net_create": [_I, _P, C.c_int32, _P, _P, C.c_float, C.c_float, _P],
    "dh_sur_net_destroy": [_P],
    "dh_sur_counts": [_I, C.POINTER(C.c_int32), C.POINTER(C.c_int32)],
    "dh_sur_net_forward_dev": [_P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P],
    "dh_sur_dose_dev": [_P, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, _P, _P,
                        C.c_int32, C.c_float, C.c_int32, _P, _P],
    "dh_sur_dose_batch_dev": [_P, C.POINTER(Grid), C.c_int32, C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, _P,
                              _P, C.c_int32, C.c_float, C.c_int32, _P, _P],
    "dh_photon_dose_surrogate_dev": [_I, _P, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P,
                                     C.c_int32, _P, _P, _P, _P, C.c_int32, C.c_float, C.c_int32, _P],
    "dh_sched_create": [_I, C.c_int32, _P],
    "dh_sched_destroy": [_P],
    "dh_sched_add_case": [_P, C.POINTER(Grid), _P, _P, _P],
    "dh_sched_add_photon": [_P, C.c_int32, C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, C.c_int32, _P,
                            C.c_int32, C.c_float, C.c_int32],
    "dh_sched_add_photon_scaled": [_P, C.c_int32, C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, C.c_int32,
                                   _P, C.c_int32, C.c_float, C.c_int32, C.c_float],
    "dh_sched_add_proton_wet": [_P, C.c_int32, C.POINTER(ProtonBeam), _P],
    "dh_sched_add_proton_dose": [_P, C.c_int32, _P, _P],
    "dh_sched_add_influence_fill": [_P, C.c_int32, _P, _P, _P, _P],
    "dh_sched_run": [_P, C.c_int32, _P],
    "dh_sched_info": [_P, _P, _P, _P, _P, _P],
    "dh_tune_set": [C.c_char_p, C.c_char_p, C.c_int32, C.c_double],
    "dh_tune_get": [C.c_char_p, C.c_char_p, C.POINTER(C.c_int32)],
    "dh_tune_count": [C.POINTER(C.c_int32)],
    "dh_tune_entry": [C.c_int32, C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, C.POINTER(C.c_int32),
                      C.POINTER(C.c_double)],
    "dh_tune_clear": [],
    "dh_autotune_proton_wet": [_I, C.POINTER(Grid), C.POINTER(ProtonBeam), _P, C.c_int32, C.POINTER(C.c_int32),
                               C.POINTER(C.c_int32), _P, _P, _P],
    "dh_autotune_pencil": [_P, _P, C.c_int32, C.POINTER(C.c_int32), _P, _P],
    "dh_autotune_photon": [_I, C.POINTER(Grid), C.POINTER(PhotonCP), C.POINTER(PhotonModel), _P, C.c_int32, _P,
                           C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P, _P, _P],
    "dh_autotune_conv3d": [_I, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                           C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32), _P, _P, C.POINTER(C.c_int32), _P],
    "dh_prof_enable": [_I],
    "dh_prof_reset": [],
    "dh_prof_get": [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)],
}
RESTYPES = {"dh_last_error": C.c_char_p}

_lib = None


class EngineError(RuntimeError):
    """A failure reported by libdosehip.so."""


def load():
    """Load libdosehip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"dosecuda_amd: {LIB_PATH} is not built; run `python -m dosecuda_amd.build` "
            "(or __graft_entry__.build()).  There is no CPU fallback.")
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = RESTYPES.get(name, C.c_int)
    _lib = L
    return L


def last_error() -> str:
    msg = load().dh_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


def check(rc: int) -> None:
    """Map a DH_E_* code to the exception the reference raises
    (dosemodule.cu:189-197): RuntimeError("CUDA error: ...") for runtime
    failures, MemoryError for host allocation failures, ValueError for bad
    arguments."""
    if rc == DH_OK:
        return
    msg = last_error()
    if rc == DH_E_NOMEM:
        raise MemoryError("Not enough host memory")
    if rc == DH_E_INVALID:
        raise ValueError(msg)
    raise RuntimeError("CUDA error: %s" % msg)


def ptr(a) -> C.c_void_p:
    """Address of a contiguous numpy array or a torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return C.c_void_p(a.ctypes.data)
    return C.c_void_p(a.data_ptr())


def grid(shape, spacing, origin) -> Grid:
    g = Grid()
    g.nz, g.ny, g.nx = (int(v) for v in shape)
    g.spacing = float(np.float32(spacing))
    o = np.asarray(origin, dtype=np.float32).reshape(-1)
    for a in range(3):
        g.origin[a] = float(o[a])
    return g


def proton_beam(iso, gantry, couch, vsadx, vsady) -> ProtonBeam:
    b = ProtonBeam()
    c = np.asarray(iso, dtype=np.float32).reshape(-1)
    for a in range(3):
        b.iso[a] = float(c[a])
    b.gantry_angle, b.couch_angle = float(gantry), float(couch)
    b.vsadx, b.vsady = float(vsadx), float(vsady)
    return b


def proton_lut(dvp, depths, sigmas, idds, keep: list) -> ProtonLut:
    """Beam-model tables read semantically: row e, column c; only the first
    400 columns are used (dosemodule.cu:345-349)."""
    arrs = []
    for a in (dvp, depths, sigmas, idds):
        arrs.append(np.ascontiguousarray(a, dtype=np.float32))
    dvp, depths, sigmas, idds = arrs
    n_e = depths.shape[0]
    if dvp.shape[0] < n_e or dvp.shape[1] < 5:
        raise ValueError("divergence_params must be [n_energies, 5]")
    dvp = np.ascontiguousarray(dvp[:, :5])
    cols = min(depths.shape[1], sigmas.shape[1], idds.shape[1])
    if sigmas.shape[0] < n_e or idds.shape[0] < n_e:
        raise ValueError("lut_sigmas and lut_idds need one row per energy of lut_depths")
    if n_e and cols < LUT_LENGTH:
        raise ValueError("lookup tables need at least 400 depth samples per energy")
    depths, sigmas, idds = (np.ascontiguousarray(a[:n_e, :LUT_LENGTH]) for a in (depths, sigmas, idds))
    keep.extend([dvp, depths, sigmas, idds])
    L = ProtonLut()
    L.n_energies, L.row_stride = n_e, LUT_LENGTH
    L.divergence_params, L.depths = dvp.ctypes.data, depths.ctypes.data
    L.sigmas, L.idds = sigmas.ctypes.data, idds.ctypes.data
    return L


def photon_cp(iso, mu, ga, ca, ta) -> PhotonCP:
    cp = PhotonCP()
    c = np.asarray(iso, dtype=np.float32).reshape(-1)
    for a in range(3):
        cp.iso[a] = float(c[a])
    cp.mu, cp.ga, cp.ca, cp.ta = float(mu), float(ga), float(ca), float(ta)
    return cp


def photon_model(values: dict, arrays: dict, kernel, keep: list) -> PhotonModel:
    """values: the 12 model floats + has_x/yjaws; arrays: the six 1-D tables;
    kernel: [6][6] rows = zenith angle, columns theta, Am, am, Bm, bm,
    ray_length (the semantic layout of kernel.csv)."""
    m = PhotonModel()
    for k in ("mu_calibration", "primary_source_distance", "scatter_source_distance",
              "primary_source_size", "scatter_source_size", "mlc_distance",
              "scatter_source_weight", "electron_attenuation", "electron_source_weight",
              "electron_fitted_dmax", "jaw_transmission", "mlc_transmission"):
        setattr(m, k, float(values[k]))
    m.has_xjaws = int(bool(values.get("has_xjaws", True)))
    m.has_yjaws = int(bool(values.get("has_yjaws", True)))
    a = {k: np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1)) for k, v in arrays.items()}
    keep.extend(a.values())
    m.n_profile = int(a["profile_radius"].size)
    for k in ("profile_intensities", "profile_softening"):
        if a[k].size < m.n_profile:
            raise ValueError("'%s' is shorter than 'profile_radius'" % k)
    m.profile_radius = a["profile_radius"].ctypes.data
    m.profile_intensities = a["profile_intensities"].ctypes.data
    m.profile_softening = a["profile_softening"].ctypes.data
    m.n_spectral = int(a["spectrum_attenuation_coefficients"].size)
    for k in ("spectrum_primary_weights", "spectrum_scatter_weights"):
        if a[k].size < m.n_spectral:
            raise ValueError("'%s' is shorter than 'spectrum_attenuation_coefficients'" % k)
    m.spectrum_attenuation_coefficients = a["spectrum_attenuation_coefficients"].ctypes.data
    m.spectrum_primary_weights = a["spectrum_primary_weights"].ctypes.data
    m.spectrum_scatter_weights = a["spectrum_scatter_weights"].ctypes.data
    k = np.asarray(kernel, dtype=np.float32)
    if k.ndim != 2 or k.shape[0] < 6 or k.shape[1] < 6:
        raise ValueError("'kernel' must hold 6 rows of (theta, Am, am, Bm, bm, ray_length)")
    flat = np.ascontiguousarray(k[:6, :6]).reshape(-1)
    for i in range(36):
        m.kernel[i] = float(flat[i])
    return m
