"""ctypes front-end of the FP64 host oracle (oracle/dose_oracle.c).

TEST INFRASTRUCTURE ONLY: the engine under dosecuda_amd/ never imports this
module.  Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of bench.py use it, to check (and time beside) the engine, never in
place of it.

The pipeline helpers chain the stages through float32 exactly where the
reference hands float32 buffers from one kernel to the next
(DoseCUDA/dose_kernels/IMPTClasses.cu:274-340, IMRTClasses.cu:353-405):
raw WET -> float32 -> smoothing -> float32 -> pencil beam; and
raw WET -> float32 -> TERMA/electron -> float32 -> CCCS.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

LUT_LENGTH = 400
AIR_DENSITY = np.float32(0.0012)


class Geom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("s", C.c_float), ("inv_s", C.c_float),
                ("iso", C.c_float * 3), ("src", C.c_float * 3),
                ("singa", C.c_float), ("cosga", C.c_float),
                ("sinta", C.c_float), ("costa", C.c_float),
                ("sinca", C.c_float), ("cosca", C.c_float)]


class Spot(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("mu", C.c_float), ("id", C.c_int)]


class Layer(C.Structure):
    _fields_ = [("start", C.c_int), ("n", C.c_int), ("id", C.c_int),
                ("r80", C.c_float), ("energy", C.c_float)]


class PhotonModel(C.Structure):
    _fields_ = [("n_profile", C.c_int), ("prof_r", C.c_void_p), ("prof_i", C.c_void_p),
                ("prof_s", C.c_void_p), ("n_spec", C.c_int), ("mu_e", C.c_void_p),
                ("pw", C.c_void_p), ("sw", C.c_void_p),
                ("mu_cal", C.c_double), ("sad", C.c_double), ("ssd_s", C.c_double),
                ("src_p", C.c_double), ("src_s", C.c_double), ("d_mlc", C.c_double),
                ("ssw", C.c_double), ("el_att", C.c_double), ("esw", C.c_double),
                ("el_dmax", C.c_double), ("kernel", C.c_float * 36)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


_lib = None


def load(path: str):
    """Open an oracle library at `path` and declare its signatures."""
    L = C.CDLL(path)
    P = C.c_void_p
    L.orc_geom_init.argtypes = [C.POINTER(Geom), C.c_int, C.c_int, C.c_int, C.c_float,
                                P, P, C.c_double, C.c_double, C.c_double, C.c_float]
    L.orc_sample.argtypes = [P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_double, C.c_double]
    L.orc_sample.restype = C.c_double
    L.orc_raytrace.argtypes = [C.POINTER(Geom), P, C.c_double, P, P]
    L.orc_raytrace_m.argtypes = [C.POINTER(Geom), P, C.c_double, P, P, P]
    L.orc_terma_m.argtypes = [C.POINTER(Geom), C.POINTER(PhotonModel), P, C.c_int, P, P, P, P]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_smooth.argtypes = [C.POINTER(Geom), P, P]
    L.orc_make_layers.argtypes = [P, C.c_int, C.c_int, P, P, P]
    L.orc_make_layers.restype = C.c_int
    L.orc_pencil.argtypes = [C.POINTER(Geom), C.c_float, C.c_float, P, P, P, C.c_int,
                             P, P, P, P, C.c_int, P]
    L.orc_pencil.restype = C.c_int64
    L.orc_terma.argtypes = [C.POINTER(Geom), C.POINTER(PhotonModel), P, C.c_int, P, P, P]
    L.orc_ccc.argtypes = [C.POINTER(Geom), C.POINTER(PhotonModel), C.c_double, P, P, P, P, P]
    L.orc_ccc.restype = C.c_int64
    L.orc_cone_steps.argtypes = [C.c_float, C.c_float]
    L.orc_cone_steps.restype = C.c_int
    L.orc_img_to_head.argtypes = [C.POINTER(Geom), C.c_int, P, P]
    L.orc_head_to_img.argtypes = [C.POINTER(Geom), C.c_int, P, P]
    L.orc_kernel_tilt.argtypes = [C.POINTER(Geom), P, P]
    L.orc_cone_dir.argtypes = [C.POINTER(Geom), P, C.c_int, C.c_int, P, P]
    return L


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = load(_LIB_PATH)
    return _lib


class using:
    """Context manager: route every helper of this module through the oracle
    library at `path` (a variant built from a modified dose_oracle.c, for the
    tests that check that the pins catch planted mistakes)."""

    def __init__(self, path: str):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = load(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def geom(shape, spacing, origin, iso, gantry, couch, collimator, src_dist) -> Geom:
    """Beam/grid set-up of dosemodule.cu:143-172 and CudaClasses.cu:17-51."""
    g = Geom()
    o = _f32(np.asarray(origin).reshape(3))
    c = _f32(np.asarray(iso).reshape(3))
    nz, ny, nx = (int(v) for v in shape)
    lib().orc_geom_init(C.byref(g), nz, ny, nx, C.c_float(float(spacing)), _p(o), _p(c),
                        float(gantry), float(couch), float(collimator),
                        C.c_float(float(src_dist)))
    return g


def img_to_head(g: Geom, p, fp32=False):
    """Image -> head (beam's-eye) coordinates of one point."""
    a = np.ascontiguousarray(p, np.float64).reshape(3)
    out = np.empty(3, np.float64)
    lib().orc_img_to_head(C.byref(g), int(bool(fp32)), _p(a), _p(out))
    return out


def head_to_img(g: Geom, h, fp32=False):
    a = np.ascontiguousarray(h, np.float64).reshape(3)
    out = np.empty(3, np.float64)
    lib().orc_head_to_img(C.byref(g), int(bool(fp32)), _p(a), _p(out))
    return out


def kernel_tilt(g: Geom, p, d):
    """kernelTilt (IMRTClasses.cu:173-203) of image-frame vector d at point p."""
    a = np.ascontiguousarray(p, np.float64).reshape(3)
    out = np.ascontiguousarray(d, np.float64).reshape(3).copy()
    lib().orc_kernel_tilt(C.byref(g), _p(a), _p(out))
    return out


def cone_dir(g: Geom, kernel, r, az, p):
    """Image-frame direction of cone (zenith row r, azimuth az) at point p."""
    k = np.ascontiguousarray(np.asarray(kernel, np.float32)[:6, :6]).reshape(36)
    a = np.ascontiguousarray(p, np.float64).reshape(3)
    out = np.empty(3, np.float64)
    lib().orc_cone_dir(C.byref(g), _p(k), int(r), int(az), _p(a), _p(out))
    return out


def sample(vol, t, border=0.0):
    vol = _f32(vol)
    nz, ny, nx = vol.shape
    return lib().orc_sample(_p(vol), nx, ny, nz, float(t[0]), float(t[1]), float(t[2]),
                            float(border))


def raytrace(g: Geom, rho, border=0.0, mask=None):
    rho = _f32(rho)
    out = np.empty(rho.shape, np.float64)
    ns = np.empty(rho.shape, np.int32)
    m8 = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_raytrace_m(C.byref(g), _p(rho), float(border), _p(out), _p(ns), _p(m8))
    return out, ns


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's loops."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def smooth(g: Geom, wet32):
    wet32 = _f32(wet32)
    out = np.empty(wet32.shape, np.float64)
    lib().orc_smooth(C.byref(g), _p(wet32), _p(out))
    return out


def make_layers(spot_rows, n_energies, dvp):
    spot_rows = _f32(np.asarray(spot_rows).reshape(-1, 4))
    n = spot_rows.shape[0]
    spots = (Spot * max(n, 1))()
    layers = (Layer * max(n_energies, 1))()
    dvp = _f32(dvp)
    nl = lib().orc_make_layers(_p(spot_rows), n, int(n_energies), _p(dvp),
                               C.cast(spots, C.c_void_p), C.cast(layers, C.c_void_p))
    return spots, layers, nl


def pencil(g: Geom, vsadx, vsady, wet32, spot_rows, dvp, lut_d, lut_s, lut_i):
    """Pencil-beam dose (FP64) and the number of voxel-spot pairs evaluated."""
    wet32 = _f32(wet32)
    dvp = _f32(dvp)
    lut_d, lut_s, lut_i = _f32(lut_d), _f32(lut_s), _f32(lut_i)
    n_e = lut_d.shape[0]
    spots, layers, nl = make_layers(spot_rows, n_e, dvp)
    out = np.empty(wet32.shape, np.float64)
    pairs = lib().orc_pencil(C.byref(g), C.c_float(float(np.float32(vsadx))),
                             C.c_float(float(np.float32(vsady))), _p(wet32),
                             C.cast(spots, C.c_void_p), C.cast(layers, C.c_void_p), nl,
                             _p(dvp), _p(lut_d), _p(lut_s), _p(lut_i), lut_d.shape[1], _p(out))
    return out, int(pairs)


def photon_model(m) -> tuple:
    """Pack an object with the reference's photon-model attribute names
    (dosemodule.cu:400-437) into the oracle's struct; the kernel table is read
    semantically as [zenith row][theta, Am, am, Bm, bm, ray_length]."""
    keep = []

    def arr(name):
        a = _f32(getattr(m, name)).reshape(-1)
        keep.append(a)
        return a

    pr, pi, ps = arr("profile_radius"), arr("profile_intensities"), arr("profile_softening")
    mu_e, pw, sw = (arr("spectrum_attenuation_coefficients"), arr("spectrum_primary_weights"),
                    arr("spectrum_scatter_weights"))
    pm = PhotonModel()
    pm.n_profile = pr.size
    pm.prof_r, pm.prof_i, pm.prof_s = pr.ctypes.data, pi.ctypes.data, ps.ctypes.data
    pm.n_spec = mu_e.size
    pm.mu_e, pm.pw, pm.sw = mu_e.ctypes.data, pw.ctypes.data, sw.ctypes.data
    # dosemodule.cu:501-515 stores these doubles into float fields
    f = lambda n: float(np.float32(float(getattr(m, n))))
    pm.mu_cal = f("mu_calibration")
    pm.sad = f("primary_source_distance")
    pm.ssd_s = f("scatter_source_distance")
    pm.src_p = f("primary_source_size")
    pm.src_s = f("scatter_source_size")
    pm.d_mlc = f("mlc_distance")
    pm.ssw = f("scatter_source_weight")
    pm.el_att = f("electron_attenuation")
    pm.esw = f("electron_source_weight")
    pm.el_dmax = f("electron_fitted_dmax")
    k = np.asarray(m.kernel, dtype=np.float32)[:6, :6]
    for r in range(6):
        for c in range(6):
            pm.kernel[6 * r + c] = float(k[r, c])
    return pm, keep


def terma(g: Geom, model, mlc_rows, wet32, mask=None):
    pm, keep = photon_model(model)
    mlc = _f32(np.asarray(mlc_rows)[:, :4])
    wet32 = _f32(wet32)
    t = np.empty(wet32.shape, np.float64)
    e = np.empty(wet32.shape, np.float64)
    m8 = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    lib().orc_terma_m(C.byref(g), C.byref(pm), _p(mlc), mlc.shape[0], _p(wet32), _p(t), _p(e), _p(m8))
    del keep
    return t, e


def ccc(g: Geom, model, mu, terma32, rho32, electron32, mask=None):
    pm, keep = photon_model(model)
    terma32, rho32, electron32 = _f32(terma32), _f32(rho32), _f32(electron32)
    out = np.empty(terma32.shape, np.float64)
    m8 = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
    n = lib().orc_ccc(C.byref(g), C.byref(pm), float(np.float32(float(mu))), _p(terma32),
                      _p(rho32), _p(electron32), _p(out), _p(m8))
    del keep
    return out, int(n)


def cone_steps(L0, spacing_mm) -> int:
    return int(lib().orc_cone_steps(C.c_float(float(L0)), C.c_float(float(spacing_mm))))


# ---------------------------------------------------------------------------
# Pipelines mirroring the three reference entry points.
# ---------------------------------------------------------------------------

def proton_wet(rlsp, spacing, origin, iso, gantry, couch, vsadx, vsady):
    """proton_raytrace_cuda (IMPTClasses.cu:274-299): returns (raw f64,
    smoothed f64, smoothed-input raw f32)."""
    D = (np.float32(vsadx) + np.float32(vsady)) / np.float32(2.0)
    g = geom(rlsp.shape, spacing, origin, iso, gantry, couch, 0.0, D)
    raw, _ = raytrace(g, rlsp, 0.0)
    raw32 = raw.astype(np.float32)
    sm = smooth(g, raw32)
    return raw, sm, raw32


def proton_dose(wet32, spacing, origin, iso, gantry, couch, vsadx, vsady, spot_rows, dvp,
                lut_d, lut_s, lut_i):
    """proton_spot_cuda (IMPTClasses.cu:301-340)."""
    D = (np.float32(vsadx) + np.float32(vsady)) / np.float32(2.0)
    g = geom(wet32.shape, spacing, origin, iso, gantry, couch, 0.0, D)
    return pencil(g, vsadx, vsady, wet32, spot_rows, dvp, lut_d, lut_s, lut_i)


def photon_dose(density, spacing, origin, iso, mu, ga, ca, ta, model, mlc_rows,
                return_stages=False):
    """photon_dose_cuda (IMRTClasses.cu:353-405)."""
    sad = np.float32(float(model.primary_source_distance))
    g = geom(density.shape, spacing, origin, iso, ga, ta, ca, sad)
    raw, _ = raytrace(g, density, float(AIR_DENSITY))
    raw32 = raw.astype(np.float32)
    t, e = terma(g, model, mlc_rows, raw32)
    t32, e32 = t.astype(np.float32), e.astype(np.float32)
    dose, n = ccc(g, model, mu, t32, _f32(density), e32)
    if return_stages:
        return dose, dict(wet=raw, wet32=raw32, terma=t, electron=e, terma32=t32,
                          electron32=e32, samples=n, geom=g)
    return dose
