"""Drop-in replacement of the reference's ``dose_kernels`` extension module.

Same three functions, positional arguments, attribute contract, dtype
checks, error types/messages and return shapes as the CPython extension in
DoseCUDA/dose_kernels/dosemodule.cu:104-595, backed by libdosehip.so (HIP,
gfx950).  The reference plan code (``plan_impt.py``, ``plan_imrt.py``) does
``import dose_kernels``; putting this module on ``sys.path`` under that name
(or ``sys.modules["dose_kernels"] = dosecuda_amd.dose_kernels``) is the whole
integration.

Arrays are read by meaning, not by raw buffer order: ``spot_list[i, c]``,
``mlc[i, c]`` and ``kernel[row, c]``.  For the layouts the reference's own
plan code produces (C-order spot lists, F-order MLC and kernel arrays, see
plan_impt.py:267, plan_imrt.py:208,373) this is exactly what the reference
reads; for other layouts the reference would read the buffer in a different
order.
"""
from __future__ import annotations

import operator

import numpy as np

from . import _lib

__all__ = ["proton_raytrace_cuda", "proton_spot_cuda", "photon_dose_cuda"]


# ---------------------------------------------------------------------------
# attribute scraping with the reference's semantics (dosemodule.cu:14-101)
# ---------------------------------------------------------------------------
def _getarray(obj, attr: str, dim: int) -> np.ndarray:
    """pyobject_getarray: AttributeError if missing; ValueError unless a
    numpy array of the given rank and dtype float32."""
    a = getattr(obj, attr)
    if not (isinstance(a, np.ndarray) and a.ndim == dim and a.dtype == np.float32):
        raise ValueError("'%s' must be %d-dimensional and of type float." % (attr, dim))
    return a


def _getfloat(obj, attr: str) -> float:
    """pyobject_getfloat: PyFloat_AsDouble semantics (float, __float__ or
    __index__; TypeError otherwise, str included)."""
    v = getattr(obj, attr)
    if isinstance(v, float):
        return v
    t = type(v)
    if hasattr(t, "__float__"):
        return float(t.__float__(v))
    if hasattr(t, "__index__"):
        return float(operator.index(v))
    raise TypeError("must be real number, not %s" % t.__name__)


def _getbool(obj, attr: str) -> bool:
    """pyobject_getbool: PyObject_IsTrue."""
    return bool(getattr(obj, attr))


def _gpu(gpu_id) -> int:
    """PyArg_ParseTuple 'i': integers and __index__ objects only."""
    if isinstance(gpu_id, float):
        raise TypeError("'float' object cannot be interpreted as an integer")
    return operator.index(gpu_id)


def _vec3(a: np.ndarray, attr: str) -> np.ndarray:
    if a.size < 3:
        raise ValueError("'%s' must hold 3 values" % attr)
    return np.asarray(a.reshape(-1)[:3], dtype=np.float32)


def _spacing0(a: np.ndarray) -> float:
    if a.size < 1:
        raise ValueError("'spacing' must not be empty")
    return float(a.reshape(-1)[0])


# ---------------------------------------------------------------------------
# proton_raytrace_cuda(model, volume, beam, gpu_id)   dosemodule.cu:104-201
# ---------------------------------------------------------------------------
def proton_raytrace_cuda(model_instance, volume_instance, beam_instance, gpu_id):
    """Smoothed water-equivalent thickness (cm) for one proton field."""
    gpu = _gpu(gpu_id)
    vsadx = _getfloat(model_instance, "VSADX")
    vsady = _getfloat(model_instance, "VSADY")
    density = _getarray(volume_instance, "voxel_data", 3)
    spacing = _getarray(volume_instance, "spacing", 1)
    origin = _getarray(volume_instance, "origin", 1)
    iso = _getarray(beam_instance, "iso", 1)
    ga = _getfloat(beam_instance, "gantry_angle")
    ta = _getfloat(beam_instance, "couch_angle")

    rho = np.ascontiguousarray(density)
    g = _lib.grid(rho.shape, _spacing0(spacing), _vec3(origin, "origin"))
    b = _lib.proton_beam(_vec3(iso, "iso"), ga, ta, vsadx, vsady)
    out = np.empty(rho.shape, dtype=np.float32)
    L = _lib.load()
    _lib.check(L.dh_proton_raytrace(gpu, g, b, _lib.ptr(rho), _lib.ptr(out)))
    return out


# ---------------------------------------------------------------------------
# proton_spot_cuda(model, volume, wet_volume, beam, gpu_id)  dosemodule.cu:243-374
# ---------------------------------------------------------------------------
def proton_spot_cuda(model_instance, volume_instance, wet_instance, beam_instance, gpu_id):
    """Pencil-beam dose of one proton field on the WET grid."""
    gpu = _gpu(gpu_id)
    vsadx = _getfloat(model_instance, "VSADX")
    vsady = _getfloat(model_instance, "VSADY")
    dvp = _getarray(model_instance, "divergence_params", 2)
    lut_d = _getarray(model_instance, "lut_depths", 2)
    lut_s = _getarray(model_instance, "lut_sigmas", 2)
    lut_i = _getarray(model_instance, "lut_idds", 2)
    _getarray(volume_instance, "voxel_data", 3)          # validated, unused (as the reference)
    spacing = _getarray(volume_instance, "spacing", 1)
    origin = _getarray(volume_instance, "origin", 1)
    wet = _getarray(wet_instance, "voxel_data", 3)
    iso = _getarray(beam_instance, "iso", 1)
    spots = _getarray(beam_instance, "spot_list", 2)
    ga = _getfloat(beam_instance, "gantry_angle")
    ta = _getfloat(beam_instance, "couch_angle")

    if spots.shape[0] and spots.shape[1] < 4:
        raise ValueError("'spot_list' rows must hold x, y, MU and energy id")
    rows = np.ascontiguousarray(spots[:, :4], dtype=np.float32)
    w = np.ascontiguousarray(wet)
    g = _lib.grid(w.shape, _spacing0(spacing), _vec3(origin, "origin"))
    b = _lib.proton_beam(_vec3(iso, "iso"), ga, ta, vsadx, vsady)
    keep: list = []
    lut = _lib.proton_lut(dvp, lut_d, lut_s, lut_i, keep)
    out = np.empty(w.shape, dtype=np.float32)
    L = _lib.load()
    _lib.check(L.dh_proton_spot(gpu, g, b, lut, _lib.ptr(rows), rows.shape[0], _lib.ptr(w),
                                _lib.ptr(out)))
    return out


# ---------------------------------------------------------------------------
# photon_dose_cuda(model, volume, cp, gpu_id)  dosemodule.cu:377-557
# ---------------------------------------------------------------------------
_PHOTON_FLOATS = ("mu_calibration", "primary_source_distance", "scatter_source_distance",
                  "primary_source_size", "scatter_source_size", "mlc_distance",
                  "scatter_source_weight", "electron_attenuation", "electron_source_weight",
                  "electron_fitted_dmax", "jaw_transmission", "mlc_transmission")
_PHOTON_ARRAYS = ("profile_radius", "profile_intensities", "profile_softening",
                  "spectrum_attenuation_coefficients", "spectrum_primary_weights",
                  "spectrum_scatter_weights")


def read_photon_model(model_instance, keep: list):
    """Attribute scrape of the photon beam model (dosemodule.cu:386-437)."""
    vals = {k: _getfloat(model_instance, k) for k in _PHOTON_FLOATS}
    vals["has_xjaws"] = _getbool(model_instance, "has_xjaws")
    vals["has_yjaws"] = _getbool(model_instance, "has_yjaws")
    arrs = {k: _getarray(model_instance, k, 1) for k in _PHOTON_ARRAYS}
    kernel = _getarray(model_instance, "kernel", 2)
    return _lib.photon_model(vals, arrs, kernel, keep)


def photon_dose_cuda(model_instance, volume_instance, cp_instance, gpu_id):
    """CCCS dose of one photon control point (before the output factor)."""
    gpu = _gpu(gpu_id)
    keep: list = []
    vals = {k: _getfloat(model_instance, k) for k in _PHOTON_FLOATS}
    vals["has_xjaws"] = _getbool(model_instance, "has_xjaws")
    vals["has_yjaws"] = _getbool(model_instance, "has_yjaws")
    arrs = {k: _getarray(model_instance, k, 1) for k in _PHOTON_ARRAYS}
    kernel = _getarray(model_instance, "kernel", 2)
    density = _getarray(volume_instance, "voxel_data", 3)
    spacing = _getarray(volume_instance, "spacing", 1)
    origin = _getarray(volume_instance, "origin", 1)
    iso = _getarray(cp_instance, "iso", 1)
    mlc = _getarray(cp_instance, "mlc", 2)
    mu = _getfloat(cp_instance, "mu")
    ga = _getfloat(cp_instance, "ga")
    ca = _getfloat(cp_instance, "ca")
    ta = _getfloat(cp_instance, "ta")

    m = _lib.photon_model(vals, arrs, kernel, keep)
    if mlc.shape[0] and mlc.shape[1] < 4:
        raise ValueError("'mlc' rows must hold x1, x2, y_offset and y_width")
    rows = np.ascontiguousarray(mlc[:, :4], dtype=np.float32)
    rho = np.ascontiguousarray(density)
    g = _lib.grid(rho.shape, _spacing0(spacing), _vec3(origin, "origin"))
    cp = _lib.photon_cp(_vec3(iso, "iso"), mu, ga, ca, ta)
    out = np.empty(rho.shape, dtype=np.float32)
    L = _lib.load()
    _lib.check(L.dh_photon_dose(gpu, g, cp, m, _lib.ptr(rows), rows.shape[0], _lib.ptr(rho),
                                _lib.ptr(out), None, None))
    return out
