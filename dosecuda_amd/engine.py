"""Device-pointer API of the engine on torch-ROCm tensors.

torch supplies device memory and the stream; all computation is
libdosehip.so.  Every function enqueues on the given (default: torch's
current) stream and returns without synchronising.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


_SIDE: dict = {}


def _side_streams(device, n: int):
    """n side streams (default priority) of a device, created once and reused."""
    key = torch.device(device).index or 0
    have = _SIDE.setdefault(key, [])
    while len(have) < n:
        have.append(torch.cuda.Stream(device=key))
    return have[:n]


def _dev(t: torch.Tensor) -> int:
    if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError("expected a contiguous float32 tensor on the GPU")
    return t.device.index if t.device.index is not None else 0


def _g(shape, spacing, origin):
    return _lib.grid(shape, spacing, origin)


def hu_convert(hu: torch.Tensor, xp, fp, out=None, stream=None) -> torch.Tensor:
    """HU -> RLSP or density on the device: float32(np.interp(hu, xp, fp))
    bit for bit (plan_impt.py:112-122, plan_imrt.py:114-124).  ``hu`` is a
    contiguous float32 or float64 tensor on the GPU; xp / fp the machine's
    HU table (``beam_models.hu_table``)."""
    if not (hu.is_cuda and hu.is_contiguous() and hu.dtype in (torch.float32, torch.float64)):
        raise ValueError("expected a contiguous float32 or float64 tensor on the GPU")
    dev = hu.device.index if hu.device.index is not None else 0
    out = torch.empty(hu.shape, dtype=torch.float32, device=hu.device) if out is None else out
    if out.shape != hu.shape:
        raise ValueError("out must have the shape of hu")
    _dev(out)
    x = np.ascontiguousarray(np.asarray(xp, dtype=np.float64))
    y = np.ascontiguousarray(np.asarray(fp, dtype=np.float64))
    if x.ndim != 1 or x.shape != y.shape:
        raise ValueError("xp and fp must be 1-D arrays of one length")
    _lib.check(_lib.load().dh_hu_convert_dev(dev, _lib.ptr(hu), int(hu.dtype == torch.float64), hu.numel(),
                                             x.ctypes.data, y.ctypes.data, x.size, _lib.ptr(out),
                                             _stream(stream)))
    return out


def proton_wet(rlsp: torch.Tensor, spacing, origin, iso, gantry, couch, vsadx, vsady,
               out: torch.Tensor = None, raw_out: torch.Tensor = None, stream=None):
    """Ray trace + lateral smoothing; returns the smoothed WET (cm)."""
    dev = _dev(rlsp)
    out = torch.empty_like(rlsp) if out is None else out
    L = _lib.load()
    _lib.check(L.dh_proton_raytrace_dev(dev, _g(rlsp.shape, spacing, origin),
                                        _lib.proton_beam(iso, gantry, couch, vsadx, vsady),
                                        _lib.ptr(rlsp), _lib.ptr(out), _lib.ptr(raw_out),
                                        _stream(stream)))
    return out


class ProtonPlan:
    """One proton field's spots/layers/LUT rows resident on the device."""

    def __init__(self, device: int, shape, spacing, origin, iso, gantry, couch, model, spot_rows):
        self.shape = tuple(int(v) for v in shape)
        self._keep: list = []
        lut = _lib.proton_lut(model.divergence_params, model.lut_depths, model.lut_sigmas,
                              model.lut_idds, self._keep)
        rows = np.ascontiguousarray(np.asarray(spot_rows, dtype=np.float32).reshape(-1, 4))
        self._h = C.c_void_p()
        L = _lib.load()
        _lib.check(L.dh_proton_plan_create(device, _g(shape, spacing, origin),
                                           _lib.proton_beam(iso, gantry, couch, model.VSADX,
                                                            model.VSADY),
                                           lut, _lib.ptr(rows), rows.shape[0], C.byref(self._h)))
        n = C.c_int32()
        ne = int(np.asarray(model.lut_depths).shape[0])
        spots = np.zeros(max(ne, 1), np.int32)
        cut = np.zeros(max(ne, 1), np.float32)
        _lib.check(L.dh_proton_plan_layers(self._h, C.byref(n), _lib.ptr(spots), _lib.ptr(cut)))
        self.n_layers = int(n.value)
        self.layer_spots = spots[:self.n_layers].copy()
        self.layer_cut = cut[:self.n_layers].copy()

    def dose(self, wet: torch.Tensor, out: torch.Tensor = None, stream=None):
        out = torch.empty_like(wet) if out is None else out
        _lib.check(_lib.load().dh_proton_plan_dose_dev(self._h, _lib.ptr(wet), _lib.ptr(out),
                                                       _stream(stream)))
        return out

    def pairs(self, wet32: np.ndarray) -> int:
        """Voxel-spot pairs that pass the range cut 10*WET <= 1.1*R80
        (IMPTClasses.cu:222-225) for the given float32 WET."""
        w = (np.asarray(wet32, dtype=np.float32) * np.float32(10.0)).ravel()
        tot = 0
        for n, c in zip(self.layer_spots, self.layer_cut):
            tot += int(n) * int(np.count_nonzero(~(w > c)))
        return tot

    def close(self):
        if self._h:
            _lib.load().dh_proton_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class InfluenceMatrix:
    """CSR dose-influence matrix of one proton field (rows = spots at MU 1,
    columns = voxel indices, values >= rel * row peak), built on the device.

    count() synchronises once to size the arrays; fill() is stream-ordered.
    Several fields can share one CSR by passing the running row/entry offsets
    (`base`) and the shared col/val tensors."""

    def __init__(self, device: int, shape, spacing, origin, iso, gantry, couch, model, spot_rows,
                 spots_per_block=None, row_chunks=None, rel=1e-4):
        self._keep: list = []
        if spots_per_block is None or row_chunks is None:
            # the autotuner's choice for this shape and spot count, else 16 / 32
            # (what it picks for the benchmark's fields)
            ch = C.c_int32()
            key = f"{int(shape[0])},{int(shape[1])},{int(shape[2])},{len(spot_rows)}"
            _lib.check(_lib.load().dh_tune_get(b"influence", key.encode(), C.byref(ch)))
            tS, tC = (ch.value // 1000, ch.value % 1000) if ch.value > 0 else (16, 32)
            spots_per_block = tS if spots_per_block is None else spots_per_block
            row_chunks = tC if row_chunks is None else row_chunks
        lut = _lib.proton_lut(model.divergence_params, model.lut_depths, model.lut_sigmas,
                              model.lut_idds, self._keep)
        rows = np.ascontiguousarray(np.asarray(spot_rows, dtype=np.float32).reshape(-1, 4))
        self.n_rows = rows.shape[0]
        self._h = C.c_void_p()
        L = _lib.load()
        _lib.check(L.dh_influence_create(device, _g(shape, spacing, origin),
                                         _lib.proton_beam(iso, gantry, couch, model.VSADX,
                                                          model.VSADY),
                                         lut, _lib.ptr(rows), rows.shape[0], int(spots_per_block),
                                         int(row_chunks), float(rel), C.byref(self._h)))
        self.nnz = None
        self.row_ptr = None

    def wet(self, rlsp: torch.Tensor, out: torch.Tensor = None, margin_scale: float = 1.0, stream=None):
        """The field's smoothed WET where the matrix can have entries (1e30 cm
        elsewhere); count() checks the footprint and falls back to the full
        grid when a spot reaches past it, so the matrix equals the one of
        proton_wet's WET."""
        out = torch.empty_like(rlsp) if out is None else out
        _lib.check(_lib.load().dh_influence_wet(self._h, _lib.ptr(rlsp), float(margin_scale), _lib.ptr(out),
                                                _stream(stream)))
        return out

    def footprint(self):
        """(box [x0, x1, y0, y1] mm in the head frame, in use, fallbacks so far)."""
        box = np.zeros(4, np.float32)
        act, fb = C.c_int32(), C.c_int64()
        _lib.check(_lib.load().dh_influence_footprint(self._h, _lib.ptr(box), C.byref(act), C.byref(fb)))
        return box, bool(act.value), int(fb.value)

    def count(self, wet: torch.Tensor, base: int = 0, stream=None) -> int:
        nnz = C.c_int64()
        rp = np.zeros(self.n_rows + 1, np.int64)
        _lib.check(_lib.load().dh_influence_count(self._h, _lib.ptr(wet), int(base), C.byref(nnz),
                                                  _lib.ptr(rp), _stream(stream)))
        self.nnz = int(nnz.value)
        self.row_ptr = rp
        return self.nnz

    @staticmethod
    def count_many(fields, rlsp: torch.Tensor, wets, base: int = 0, margin_scale: float = 1.0,
                   stream=None):
        """wet() and count() of several fields of one device at once (the
        fields' passes overlap on library streams, one host synchronisation);
        field q's entries follow those of fields 0..q-1 from `base`.  Returns
        the per-field entry counts; each field's nnz and row_ptr are set as
        count() sets them."""
        n = len(fields)
        if len(wets) != n:
            raise ValueError("one WET buffer per field")
        hs = (C.c_void_p * max(n, 1))(*[f._h for f in fields])
        ws = (C.c_void_p * max(n, 1))(*[_lib.ptr(w) for w in wets])
        nnz = np.zeros(max(n, 1), np.int64)
        rps = [np.zeros(f.n_rows + 1, np.int64) for f in fields]
        rpp = (C.c_void_p * max(n, 1))(*[_lib.ptr(r) for r in rps])
        _lib.check(_lib.load().dh_influence_count_many(hs, n, _lib.ptr(rlsp), float(margin_scale), ws, int(base),
                                                       _lib.ptr(nnz), rpp, _stream(stream)))
        for q, f in enumerate(fields):
            f.nnz = int(nnz[q])
            f.row_ptr = rps[q]
        return [int(x) for x in nnz[:n]]

    @staticmethod
    def fill_many(fields, wets, col: torch.Tensor, val: torch.Tensor, stream=None):
        """fill() of several counted fields into one shared CSR, each on a
        side stream of its own forked from `stream` (the current stream by
        default) and joined back to it, so the fills overlap on the GPU; the
        entries are those of the fills in turn (each field writes only its own
        slots)."""
        if len(wets) != len(fields):
            raise ValueError("one WET buffer per field")
        cur = stream if stream is not None else torch.cuda.current_stream()
        side = _side_streams(cur.device, len(fields))
        for s in side:
            s.wait_stream(cur)
        for f, w, s in zip(fields, wets, side):
            f.fill(w, col, val, stream=s)
            # keep the buffers alive until the side stream has used them
            for t in (w, col, val):
                t.record_stream(s)
        for s in side:
            cur.wait_stream(s)

    def fill(self, wet: torch.Tensor, col: torch.Tensor, val: torch.Tensor, stream=None):
        _lib.check(_lib.load().dh_influence_fill(self._h, _lib.ptr(wet), _lib.ptr(col),
                                                 _lib.ptr(val), _stream(stream)))

    def info(self):
        """(caller's spot row of each matrix row, row peak)."""
        n = C.c_int32()
        spot = np.zeros(max(self.n_rows, 1), np.int32)
        peak = np.zeros(max(self.n_rows, 1), np.float32)
        _lib.check(_lib.load().dh_influence_info(self._h, C.byref(n), _lib.ptr(spot), _lib.ptr(peak)))
        return spot[:self.n_rows], peak[:self.n_rows]

    def stats(self):
        """Sweep counters {pass: (lane visits, in-range visits, chunks, rows)}
        of the last passes and the per-row truncation radius (mm)."""
        st = np.zeros(12, np.uint64)
        rad = np.zeros(max(self.n_rows, 1), np.float32)
        _lib.check(_lib.load().dh_influence_stats(self._h, _lib.ptr(st), _lib.ptr(rad)))
        return ({"peak": tuple(int(x) for x in st[0:4]), "count": tuple(int(x) for x in st[4:8]),
                 "fill": tuple(int(x) for x in st[8:12])}, rad[:self.n_rows])

    def build(self, wet: torch.Tensor, stream=None):
        """count + fill into fresh tensors; returns (row_ptr int64, col int32, val float32)."""
        nnz = self.count(wet, 0, stream)
        col = torch.empty(max(nnz, 1), dtype=torch.int32, device=wet.device)
        val = torch.empty(max(nnz, 1), dtype=torch.float32, device=wet.device)
        self.fill(wet, col, val, stream)
        return torch.from_numpy(self.row_ptr).to(wet.device), col[:nnz], val[:nnz]

    def close(self):
        if self._h:
            _lib.load().dh_influence_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def photon_args(model, keep):
    vals = {k: getattr(model, k) for k in (
        "mu_calibration", "primary_source_distance", "scatter_source_distance",
        "primary_source_size", "scatter_source_size", "mlc_distance", "scatter_source_weight",
        "electron_attenuation", "electron_source_weight", "electron_fitted_dmax",
        "jaw_transmission", "mlc_transmission", "has_xjaws", "has_yjaws")}
    arrs = {k: getattr(model, k) for k in (
        "profile_radius", "profile_intensities", "profile_softening",
        "spectrum_attenuation_coefficients", "spectrum_primary_weights",
        "spectrum_scatter_weights")}
    return _lib.photon_model(vals, arrs, model.kernel, keep)


class PhotonParams:
    """The C-ABI struct of one photon beam model, built once and reused for
    every control point of a beam.  With ``validate`` the attributes are read
    with the drop-in's checks (dosemodule.cu:386-437), so a malformed model
    raises the same error as ``dose_kernels.photon_dose_cuda``."""

    def __init__(self, model, validate=True):
        self._keep: list = []
        if validate:
            from .dose_kernels import read_photon_model
            self.struct = read_photon_model(model, self._keep)
        else:
            self.struct = photon_args(model, self._keep)


def photon_dose(density: torch.Tensor, spacing, origin, iso, mu, ga, ca, ta, model, mlc,
                out=None, terma_out=None, electron_out=None, wet_out=None, stream=None):
    """Ray trace + TERMA + CCCS for one control point.  ``model`` is a beam
    model object or a PhotonParams built from one."""
    dev = _dev(density)
    out = torch.empty_like(density) if out is None else out
    keep: list = []
    m = model.struct if isinstance(model, PhotonParams) else photon_args(model, keep)
    rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
    _lib.check(_lib.load().dh_photon_dose_dev(
        dev, _g(density.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta), m,
        _lib.ptr(rows), rows.shape[0], _lib.ptr(density), _lib.ptr(out), _lib.ptr(terma_out),
        _lib.ptr(electron_out), _lib.ptr(wet_out), _stream(stream)))
    return out


def photon_terma(density: torch.Tensor, spacing, origin, iso, mu, ga, ca, ta, model, mlc, step=1.0,
                 terma_out=None, electron_out=None, stream=None):
    """TERMA and electron-contamination volumes of one control point, with
    the radiological depth from a ray trace sampled every ``step`` mm (1: the
    reference's ray trace), or with ``step`` 0 from fan-line integrals on a
    beam's-eye-view grid (the checked surrogate path's depth)."""
    dev = _dev(density)
    terma_out = torch.empty_like(density) if terma_out is None else terma_out
    electron_out = torch.empty_like(density) if electron_out is None else electron_out
    keep: list = []
    m = model.struct if isinstance(model, PhotonParams) else photon_args(model, keep)
    rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
    _lib.check(_lib.load().dh_photon_terma_dev(
        dev, _g(density.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta), m,
        _lib.ptr(rows), rows.shape[0], _lib.ptr(density), float(step), _lib.ptr(terma_out),
        _lib.ptr(electron_out), _stream(stream)))
    return terma_out, electron_out


def photon_terma_batch(density: torch.Tensor, spacing, origin, cps, model, mlcs, step=1.0, terma_outs=None,
                       electron_outs=None, stream=None):
    """photon_terma of several control points of one grid and density
    (cps: (iso, mu, ga, ca, ta) per point, mlcs: their apertures) in one call
    of the library: the density is padded once and the points' depth and
    TERMA launches alternate between the stream and a side stream.  Each
    output equals photon_terma of that point bit for bit."""
    dev = _dev(density)
    n = len(cps)
    if n == 0 or len(mlcs) != n:
        raise ValueError("need one aperture per control point")
    terma_outs = [torch.empty_like(density) for _ in cps] if terma_outs is None else list(terma_outs)
    electron_outs = [torch.empty_like(density) for _ in cps] if electron_outs is None else list(electron_outs)
    if len(terma_outs) != n or len(electron_outs) != n:
        raise ValueError("need one TERMA and one electron volume per control point")
    for v in (*terma_outs, *electron_outs):
        if v.shape != density.shape or v.dtype != torch.float32 or not v.is_contiguous():
            raise ValueError("volumes must be contiguous float32 of the density's shape")
    keep: list = []
    m = model.struct if isinstance(model, PhotonParams) else photon_args(model, keep)
    rows = [np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4]) for mlc in mlcs]
    arr = (_lib.PhotonCP * n)(*[_lib.photon_cp(*cp) for cp in cps])
    P = lambda vs: (C.c_void_p * n)(*[v.data_ptr() for v in vs])
    R = (C.c_void_p * n)(*[r.ctypes.data for r in rows])
    npairs = (C.c_int32 * n)(*[r.shape[0] for r in rows])
    _lib.check(_lib.load().dh_photon_terma_batch_dev(
        dev, _g(density.shape, spacing, origin), n, arr, m, R, npairs, _lib.ptr(density), float(step),
        P(terma_outs), P(electron_outs), _stream(stream)))
    return terma_outs, electron_outs


# ---- single stages (stage-wise parity) -----------------------------------
def stage_raytrace(vol: torch.Tensor, spacing, origin, iso, gantry, couch, src_dist, border,
                   out=None, stream=None):
    out = torch.empty_like(vol) if out is None else out
    c = np.asarray(iso, dtype=np.float32).reshape(3)
    _lib.check(_lib.load().dh_stage_raytrace_dev(
        _dev(vol), _g(vol.shape, spacing, origin), _lib.ptr(c), float(gantry), float(couch),
        float(src_dist), float(border), _lib.ptr(vol), _lib.ptr(out), _stream(stream)))
    return out


def stage_smooth(wet_raw: torch.Tensor, spacing, origin, iso, gantry, couch, vsadx, vsady,
                 out=None, stream=None):
    out = torch.empty_like(wet_raw) if out is None else out
    _lib.check(_lib.load().dh_stage_smooth_dev(
        _dev(wet_raw), _g(wet_raw.shape, spacing, origin),
        _lib.proton_beam(iso, gantry, couch, vsadx, vsady), _lib.ptr(wet_raw), _lib.ptr(out),
        _stream(stream)))
    return out


def stage_terma(wet: torch.Tensor, spacing, origin, iso, mu, ga, ca, ta, model, mlc,
                stream=None):
    terma, el = torch.empty_like(wet), torch.empty_like(wet)
    keep: list = []
    m = photon_args(model, keep)
    rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
    _lib.check(_lib.load().dh_stage_terma_dev(
        _dev(wet), _g(wet.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta), m,
        _lib.ptr(rows), rows.shape[0], _lib.ptr(wet), _lib.ptr(terma), _lib.ptr(el),
        _stream(stream)))
    return terma, el


def stage_ccc(terma: torch.Tensor, density: torch.Tensor, electron: torch.Tensor, spacing,
              origin, iso, mu, ga, ca, ta, model, out=None, stream=None):
    out = torch.empty_like(terma) if out is None else out
    keep: list = []
    m = photon_args(model, keep)
    _lib.check(_lib.load().dh_stage_ccc_dev(
        _dev(terma), _g(terma.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta), m,
        _lib.ptr(terma), _lib.ptr(density), _lib.ptr(electron), _lib.ptr(out), _stream(stream)))
    return out


# ---- launch timing -------------------------------------------------------
def prof_enable(on=True):
    _lib.load().dh_prof_enable(1 if on else 0)


def prof_reset():
    _lib.load().dh_prof_reset()


def prof_get(kernel: str):
    ms, n = C.c_double(), C.c_int64()
    _lib.check(_lib.load().dh_prof_get(kernel.encode(), C.byref(ms), C.byref(n)))
    return float(ms.value), int(n.value)
