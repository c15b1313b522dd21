"""Autotuner front end.  The library's tuners (dh_autotune_*) time every
variant of a kernel on the caller's data, check that the variants agree and
store the fastest in the library's tuning table, which the launchers consult
per shape.  This module runs them, tunes the influence matrix's spots per
workgroup and row chunks (which are handle parameters), and keeps every
choice in a JSON cache per GPU architecture.

No reference counterpart: the reference's launch shapes are compile-time
constants (CudaClasses.cuh:14).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np
import torch

from . import _lib, engine

# voxel tiles of the ray trace, smoothing and pencil launches (the library's
# choice codes); "/y": the same bricks numbered y fastest
TILES = ["linear", "64x4x1", "32x8x1", "16x16x1", "32x4x2", "16x4x4", "8x8x4", "64x2x2",
         "16x16x1/y", "16x4x4/y", "8x8x4/y", "32x4x2/y"]


def arch(device=0) -> str:
    try:
        return torch.cuda.get_device_properties(device).gcnArchName.split(":")[0]
    except Exception:
        return "gfx950"


def cache_path(device=0) -> str:
    p = os.environ.get("DOSECUDA_AMD_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "dosecuda_amd", f"tune_{arch(device)}.json")


def entries():
    L = _lib.load()
    n = C.c_int32()
    _lib.check(L.dh_tune_count(C.byref(n)))
    out = []
    for i in range(n.value):
        k, key = C.create_string_buffer(64), C.create_string_buffer(256)
        ch, ms = C.c_int32(), C.c_double()
        _lib.check(L.dh_tune_entry(i, k, 64, key, 256, C.byref(ch), C.byref(ms)))
        out.append(dict(kernel=k.value.decode(), key=key.value.decode(), choice=int(ch.value), ms=float(ms.value)))
    return out


def get(kernel: str, key: str):
    ch = C.c_int32()
    _lib.check(_lib.load().dh_tune_get(kernel.encode(), key.encode(), C.byref(ch)))
    return None if ch.value < 0 else int(ch.value)


def set_choice(kernel: str, key: str, choice: int, ms: float = 0.0):
    _lib.check(_lib.load().dh_tune_set(kernel.encode(), key.encode(), int(choice), float(ms)))


def clear():
    _lib.check(_lib.load().dh_tune_clear())


def save(path=None, device=0):
    path = path or cache_path(device)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    doc = dict(arch=arch(device), saved=time.strftime("%Y-%m-%dT%H:%M:%S"), entries=entries())
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    return path


def load(path=None, device=0) -> int:
    """Loads cached choices for this architecture into the library's table;
    returns how many were applied (0 when there is no cache)."""
    path = path or cache_path(device)
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        doc = json.load(f)
    if doc.get("arch") != arch(device):
        return 0
    for e in doc.get("entries", []):
        set_choice(e["kernel"], e["key"], e["choice"], e.get("ms", 0.0))
    return len(doc.get("entries", []))


def tune_proton_wet(rlsp: torch.Tensor, spacing, origin, iso, gantry, couch, model, reps=5, stream=None):
    br, bs = C.c_int32(), C.c_int32()
    mr, ms = np.zeros(len(TILES), np.float32), np.zeros(len(TILES), np.float32)
    _lib.check(_lib.load().dh_autotune_proton_wet(
        engine._dev(rlsp), _lib.grid(rlsp.shape, spacing, origin),
        _lib.proton_beam(iso, gantry, couch, model.VSADX, model.VSADY), _lib.ptr(rlsp), int(reps), C.byref(br),
        C.byref(bs), _lib.ptr(mr), _lib.ptr(ms), engine._stream(stream)))
    return dict(raytrace=TILES[br.value], raytrace_ms=mr.tolist(), smooth=TILES[bs.value], smooth_ms=ms.tolist())


def tune_photon(density: torch.Tensor, spacing, origin, iso, mu, ga, ca, ta, model, mlc, reps=5, stream=None):
    """Ray-trace tile and cone-superposition layout of one photon control
    point (the choices apply to every control point of that shape and beam
    axis)."""
    keep: list = []
    m = engine.photon_args(model, keep)
    rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
    br, bl = C.c_int32(), C.c_int32()
    mr, mc = np.zeros(len(TILES), np.float32), np.zeros(2, np.float32)
    _lib.check(_lib.load().dh_autotune_photon(
        engine._dev(density), _lib.grid(density.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta), m,
        _lib.ptr(rows), rows.shape[0], _lib.ptr(density), int(reps), C.byref(br), C.byref(bl), _lib.ptr(mr),
        _lib.ptr(mc), engine._stream(stream)))
    return dict(raytrace=TILES[br.value], raytrace_ms=mr.tolist(), ccc_layout=["pairs", "quads"][bl.value],
                ccc_ms=mc.tolist())


def tune_pencil(plan: "engine.ProtonPlan", wet: torch.Tensor, reps=5, stream=None):
    b = C.c_int32()
    ms = np.zeros(2 * len(TILES), np.float32)
    _lib.check(_lib.load().dh_autotune_pencil(plan._h, _lib.ptr(wet), int(reps), C.byref(b), _lib.ptr(ms),
                                              engine._stream(stream)))
    c = int(b.value)
    return dict(tile=TILES[c % len(TILES)], lut_in_lds=c < len(TILES), choice=c, ms=ms.tolist())


def tune_conv3d(n, d, h, w, ca, cb, up, cout, ksize, reps=5, device=0, stream=None):
    best, nt = C.c_int32(), C.c_int32()
    codes = np.zeros(24, np.int32)
    ms = np.zeros(24, np.float32)
    _lib.check(_lib.load().dh_autotune_conv3d(device, n, d, h, w, ca, cb, 1 if up else 0, cout, ksize, int(reps),
                                              C.byref(best), _lib.ptr(codes), _lib.ptr(ms), C.byref(nt),
                                              engine._stream(stream)))
    k = int(nt.value)
    return dict(best=int(best.value), codes=codes[:k].tolist(), ms=ms[:k].tolist())


def surrogate_layer_shapes(net, cells):
    """(n, d, h, w, ca, cb, up, cout, ksize) of every convolution of the
    surrogate network's forward pass over `cells` = (cz, cy, cx), batch 1."""
    cz, cy, cx = cells
    a, b, c, d = net.widths
    from .surrogate import IN_CH
    L = [(cz, cy, cx), (cz // 2, cy // 2, cx // 2), (cz // 4, cy // 4, cx // 4), (cz // 8, cy // 8, cx // 8)]
    s = lambda lv, ca, cb, up, co, k=3: (1, *L[lv], ca, cb, up, co, k)
    return [s(0, IN_CH, 0, 0, a), s(0, a, 0, 0, a), s(1, a, 0, 0, b), s(1, b, 0, 0, b), s(2, b, 0, 0, c),
            s(2, c, 0, 0, c), s(3, c, 0, 0, d), s(3, d, 0, 0, d), s(2, d, c, 1, c), s(2, c, 0, 0, c),
            s(1, c, b, 1, b), s(1, b, 0, 0, b), s(0, b, a, 1, a), s(0, a, 0, 0, a), s(0, a, 0, 0, 16, 1)]


def conv_key(n, d, h, w, ca, cb, up, cout, ksize):
    """The library's tuning key of a convolution shape."""
    return f"{n},{d},{h},{w},{ca},{cb},{1 if up else 0},{cout},{ksize}"


def tune_surrogate(net, cells, reps=5, batch=1):
    """Tunes every layer of the surrogate's forward pass over `cells`.  With
    batch > 1 each layer is timed over `batch` samples and the fastest
    variant is stored for that shape and for the one-sample shape too: a
    batch runs every layer with its one-sample variant (so a batched dose
    equals the single-point one bit for bit), and that variant is then the
    batch's fastest."""
    out = {}
    for shp in dict.fromkeys(surrogate_layer_shapes(net, cells)):
        r = tune_conv3d(batch, *shp[1:], reps=reps)
        if batch > 1 and r["best"] > 0:
            set_choice("conv3d", conv_key(1, *shp[1:]), r["best"])
        out[",".join(map(str, shp))] = r
    return out


def geom_key(shape, gantry, couch, src_dist):
    """The library's tuning key of a beam geometry: "nz,ny,nx,axis", axis the
    image axis of the source position's largest component (FP32, as the
    library computes it)."""
    f = np.float32
    pi = f(np.pi)
    ga = f(np.fmod(f(gantry) + f(180.0), f(360.0))) * pi / f(180.0)
    ta = f(couch) * pi / f(180.0)
    xg = -f(src_dist) * np.sin(ga, dtype=f)
    a = [abs(xg * np.cos(ta, dtype=f)), abs(f(src_dist) * np.cos(ga, dtype=f)), abs(-xg * np.sin(ta, dtype=f))]
    axis = 0 if (a[0] >= a[1] and a[0] >= a[2]) else (1 if a[1] >= a[2] else 2)
    return f"{int(shape[0])},{int(shape[1])},{int(shape[2])},{axis}"


def tune_influence_wet(rlsp: torch.Tensor, shape, spacing, origin, iso, gantry, couch, model, spots,
                       choices=(5, 6, 3, 9, 10, 8), reps=3):
    """Voxel tiles of an influence matrix's footprint ray trace and smoothing
    ("raytrace_box", "smooth_box"): times InfluenceMatrix.wet() with each
    choice, requires the identical WET of every one and keeps the fastest."""
    f = engine.InfluenceMatrix(0, shape, spacing, origin, iso, gantry, couch, model, spots)
    key = geom_key(shape, gantry, couch, (float(np.float32(model.VSADX)) + float(np.float32(model.VSADY))) / 2.0)
    out = {}
    try:
        w = torch.empty_like(rlsp)
        ref = None
        for kernel in ("raytrace_box", "smooth_box"):
            times = {}
            for c in choices:
                set_choice(kernel, key, c)
                f.wet(rlsp, out=w)
                torch.cuda.synchronize()
                if ref is None:
                    ref = w.clone()
                elif not torch.equal(ref, w):
                    raise RuntimeError(f"{kernel} tile {TILES[c]} changes the WET")
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                for _ in range(reps):
                    f.wet(rlsp, out=w)
                e1.record()
                torch.cuda.synchronize()
                times[c] = e0.elapsed_time(e1) / reps
            best = min(times, key=times.get)
            set_choice(kernel, key, best, times[best])
            out[kernel] = dict(best=TILES[best], ms={TILES[c]: t for c, t in times.items()})
    finally:
        f.close()
    return out


def tune_influence(rlsp_wet: torch.Tensor, shape, spacing, origin, iso, gantry, couch, model, spots,
                   candidates=((4, 32), (8, 16), (8, 32), (8, 64), (16, 16), (16, 32), (16, 64)), reps=3):
    """Spots per workgroup and row chunks of an influence matrix: builds the
    matrix with each pair (count + fill), requires identical CSR arrays and
    keeps the fastest fill under the key "nz,ny,nx,n_spots"."""
    res, ref, best = [], None, None
    for S, Cc in candidates:
        f = engine.InfluenceMatrix(0, shape, spacing, origin, iso, gantry, couch, model, spots,
                                   spots_per_block=S, row_chunks=Cc)
        nnz = f.count(rlsp_wet)
        col = torch.empty(nnz, dtype=torch.int32, device="cuda")
        val = torch.empty(nnz, dtype=torch.float32, device="cuda")
        f.fill(rlsp_wet, col, val)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            f.count(rlsp_wet)
            f.fill(rlsp_wet, col, val)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        if ref is None:
            ref = (col.clone(), val.clone())
        else:
            if not (torch.equal(ref[0], col) and torch.equal(ref[1], val)):
                raise RuntimeError(f"influence variants disagree at spots_per_block={S}, row_chunks={Cc}")
        res.append(dict(spots_per_block=S, row_chunks=Cc, ms=ms))
        if best is None or ms < best["ms"]:
            best = res[-1]
        f.close()
    key = f"{shape[0]},{shape[1]},{shape[2]},{len(spots)}"
    set_choice("influence", key, best["spots_per_block"] * 1000 + best["row_chunks"], best["ms"])
    return dict(best=best, all=res)
