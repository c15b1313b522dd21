"""Batching scheduler (libdosehip.so, dh_sched_*): many independent dose
cases on one GPU, spread over HIP streams with a scratch workspace each,
longest case first onto the least-loaded stream (measured job times after an
eager run), optionally captured into one HIP graph and replayed.

No reference counterpart: DoseCUDA runs one synchronous call per beam
(dosemodule.cu:95-538).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, engine


class Scheduler:
    def __init__(self, device: int = 0, streams: int = 4):
        self.device = int(device)
        h = C.c_void_p()
        _lib.check(_lib.load().dh_sched_create(self.device, int(streams), C.byref(h)))
        self._h = h
        self._keep: list = []          # tensors and handles the queued jobs point at
        self.n_streams = int(streams)

    def close(self):
        if self._h:
            _lib.load().dh_sched_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_case(self, volume: torch.Tensor, spacing, origin, dose: torch.Tensor = None) -> int:
        """A phantom: density (photon jobs) or RLSP (proton jobs) on the
        device, and the case's dose output (zeroed, then accumulated by the
        case's photon and proton-dose jobs on every run)."""
        engine._dev(volume)
        if dose is not None:
            engine._dev(dose)
            if dose.shape != volume.shape:
                raise ValueError("dose must have the volume's shape")
        cid = C.c_int32()
        _lib.check(_lib.load().dh_sched_add_case(self._h, _lib.grid(volume.shape, spacing, origin), _lib.ptr(volume),
                                                 _lib.ptr(dose), C.byref(cid)))
        self._keep += [volume, dose]
        return int(cid.value)

    def add_photon(self, case: int, iso, mu, ga, ca, ta, model, mlc, surrogate=None, stride=3, tol=0.01,
                   dilate=1, scale=1.0):
        """One photon control point; surrogate = a surrogate.NativeNet (the
        checked surrogate replaces the cone superposition) or None (plain).
        ``model`` is a beam model or an engine.PhotonParams; the dose is
        multiplied by ``scale`` (rounded to float32) before it is added to
        the case dose, so a case sums its control points' scaled doses in
        the order they were added."""
        keep: list = []
        m = model.struct if isinstance(model, engine.PhotonParams) else engine.photon_args(model, keep)
        rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
        _lib.check(_lib.load().dh_sched_add_photon_scaled(
            self._h, int(case), _lib.photon_cp(iso, mu, ga, ca, ta), m, _lib.ptr(rows), rows.shape[0],
            surrogate.handle if surrogate is not None else None, int(stride), float(tol), int(dilate),
            float(np.float32(scale))))
        if surrogate is not None:
            self._keep.append(surrogate)

    def add_proton_wet(self, case: int, iso, gantry, couch, model, wet_out: torch.Tensor):
        engine._dev(wet_out)
        _lib.check(_lib.load().dh_sched_add_proton_wet(
            self._h, int(case), _lib.proton_beam(iso, gantry, couch, model.VSADX, model.VSADY), _lib.ptr(wet_out)))
        self._keep.append(wet_out)

    def add_proton_dose(self, case: int, plan: "engine.ProtonPlan", wet: torch.Tensor):
        _lib.check(_lib.load().dh_sched_add_proton_dose(self._h, int(case), plan._h, _lib.ptr(wet)))
        self._keep += [plan, wet]

    def add_influence_fill(self, case: int, infl: "engine.InfluenceMatrix", wet: torch.Tensor,
                           col: torch.Tensor, val: torch.Tensor):
        """The fill pass of a counted influence matrix (count() ran on the
        same WET) into col/val at the matrix's row offsets."""
        _lib.check(_lib.load().dh_sched_add_influence_fill(self._h, int(case), infl._h, _lib.ptr(wet),
                                                           _lib.ptr(col), _lib.ptr(val)))
        self._keep += [infl, wet, col, val]

    def run(self, graph: bool = False, stream=None):
        """Enqueue the batch after the work already on `stream` (default: the
        current torch stream); graph=True replays one captured HIP graph
        (captured on first use after an eager warm-up)."""
        _lib.check(_lib.load().dh_sched_run(self._h, 1 if graph else 0, engine._stream(stream)))

    def info(self):
        nc, nj, ms = C.c_int32(), C.c_int32(), C.c_double()
        _lib.check(_lib.load().dh_sched_info(self._h, C.byref(nc), C.byref(nj), C.byref(ms), None, None))
        cs = np.zeros(max(nc.value, 1), np.int32)
        jm = np.zeros(max(nj.value, 1), np.float64)
        _lib.check(_lib.load().dh_sched_info(self._h, None, None, None, _lib.ptr(cs), _lib.ptr(jm)))
        return dict(cases=int(nc.value), jobs=int(nj.value), last_ms=float(ms.value),
                    case_stream=cs[:nc.value].tolist(), job_ms=jm[:nj.value].tolist())
