"""Benchmark of the MI355X dose engine (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Headline workload (one step): the plain IMRT CCCS plan of a 128^3 water
phantom at 2.5 mm with a seeded lung ellipsoid and bone cylinder, 7 coplanar
6 MV beams (gantry 360*b/7) with a 10x10 cm MLC field, VarianTrueBeamHF
model: per beam ray trace + TERMA + 72-cone superposition; density resident
in HBM.  `value` is ms per beam.  The same line carries the plain IMPT field
(100^3 at 2.5 mm, water with a seeded 2 cm bone slab, 20x20 spots x 10
layers of 100-160 MeV, unit MU) in Gvoxel-spot/s, the roofline of the
dominant kernel from HIP events around its launches, the FP64 oracle timed
on a bounded sample on the host cores, the batched influence matrix
(160x256x256 at 2 mm, 3 fields, ~15k spots), the learned surrogate CCCS on
the same 7-beam plan (gamma against the plain CCCS, speed-up, MFMA use) and
the mixed workload of 64 seeded cases run as one HIP graph.  The launch
shapes are autotuned first (--no-tune skips it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per k_ccc launch on the 128^3 plan (mean over the 7 beams) from
# `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` runs of this bench,
# scaled by a 256 MiB copy calibration (factors 2.0 and 1.0); the commands and
# per-kernel figures are in profiles/hbm_traffic_cccs128_7beams.md
CCC_TRAFFIC_BYTES = (2.0 * 111437 + 32833) * 1024
# HBM bytes per influence fill launch (mean over the 3 fields' fills) from
# `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` runs of this bench with
# the same calibration; commands and per-kernel figures in
# profiles/kernel_stats_influence_3fields.md
INFL_FILL_TRAFFIC_BYTES = (2.0 * 639669 + 1829090) * 1024


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="approximate budget of the CPU baseline sample")
    p.add_argument("--workload", default="all",
                   choices=["all", "cccs", "impt", "influence", "surrogate", "surrogate_path", "mixed"],
                   help="surrogate_path (not part of all): only the whole checked surrogate path of the "
                        "7-beam plan as a replayed HIP graph, --steps replays, for kernel traces")
    p.add_argument("--influence-steps", type=int, default=3)
    p.add_argument("--no-tune", action="store_true",
                   help="skip the autotuners (launch shapes from the cache or the defaults)")
    p.add_argument("--mixed-cases", type=int, default=64)
    p.add_argument("--mixed-steps", type=int, default=1)
    p.add_argument("--copy-calib-mb", type=int, default=0,
                   help="after the run, copy a buffer of this many MiB once (a counter "
                        "calibration kernel of known traffic for rocprofv3 --pmc runs)")
    return p.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
def setup_cccs(torch):
    from dosecuda_amd import phantoms
    from dosecuda_amd.beam_models import PhotonBeamModel, square_field_mlc
    pm = PhotonBeamModel()
    case = phantoms.imrt_lung_bone_case(seed=1002, n=128, spacing=2.5)
    dens = phantoms.to_density(case)
    mlc = square_field_mlc(pm, 10.0, 10.0)
    d = torch.from_numpy(dens).cuda()
    return dict(pm=pm, case=case, dens=dens, d_dens=d, mlc=mlc, angles=phantoms.imrt_beams(7),
                out=torch.zeros_like(d), acc=torch.zeros_like(d))


def run_cccs(torch, st):
    from dosecuda_amd import engine
    c = st["case"]
    st["acc"].zero_()
    for ga in st["angles"]:
        engine.photon_dose(st["d_dens"], c.spacing, c.origin, c.iso, 100.0, ga, 0.0, 0.0,
                           st["pm"], st["mlc"], out=st["out"])
        st["acc"].add_(st["out"])


def setup_impt(torch):
    from dosecuda_amd import engine, phantoms
    from dosecuda_amd.beam_models import ProtonBeamModel
    m = ProtonBeamModel()
    case = phantoms.impt_slab_case(seed=1001, n=100, spacing=2.5)
    rlsp = phantoms.to_rlsp(case)
    spots, ids = phantoms.impt_spot_grid(m, seed=1001, n_side=20, pitch=5.0, n_layers=10,
                                         e_lo=100.0, e_hi=160.0)
    d = torch.from_numpy(rlsp).cuda()
    plan = engine.ProtonPlan(0, rlsp.shape, case.spacing, case.origin, case.iso, 0.0, 0.0, m, spots)
    return dict(m=m, case=case, rlsp=rlsp, d_rlsp=d, spots=spots, plan=plan,
                wet=torch.empty_like(d), dose=torch.empty_like(d))


def run_impt(torch, st):
    from dosecuda_amd import engine
    c, m = st["case"], st["m"]
    engine.proton_wet(st["d_rlsp"], c.spacing, c.origin, c.iso, 0.0, 0.0, m.VSADX, m.VSADY,
                      out=st["wet"])
    st["plan"].dose(st["wet"], out=st["dose"])


def setup_influence(torch, tune=True):
    from dosecuda_amd import autotune, engine, phantoms
    from dosecuda_amd.beam_models import ProtonBeamModel
    m = ProtonBeamModel()
    case = phantoms.spheres_case(seed=1003)
    rlsp = phantoms.to_rlsp(case)
    d = torch.from_numpy(rlsp).cuda()
    fields = phantoms.influence_fields(m, seed=1003)
    st = dict(m=m, case=case, d_rlsp=d, fields=fields, wets=[], infl=[], tuning=[])
    for g, spots in fields:
        w = engine.proton_wet(d, case.spacing, case.origin, case.iso, g, 0.0, m.VSADX, m.VSADY)
        if tune:
            # spots per workgroup and row chunks of this field (identical CSR
            # required of every candidate); the matrix below reads the choice
            t = autotune.tune_influence(w, rlsp.shape, case.spacing, case.origin, case.iso, g, 0.0,
                                        m, spots, reps=2)
            # voxel tiles of the footprint ray trace and smoothing (identical WET)
            tw = autotune.tune_influence_wet(d, rlsp.shape, case.spacing, case.origin, case.iso, g, 0.0,
                                             m, spots, reps=2)
            st["tuning"].append(dict(t["best"], raytrace_box=tw["raytrace_box"]["best"],
                                     smooth_box=tw["smooth_box"]["best"]))
        f = engine.InfluenceMatrix(0, rlsp.shape, case.spacing, case.origin, case.iso, g, 0.0, m, spots)
        st["infl"].append(f)
        st["wets"].append(f.wet(d, out=w))
    # size the shared CSR once
    base = 0
    for f, w in zip(st["infl"], st["wets"]):
        base += f.count(w, base)
    st["nnz"] = base
    st["col"] = torch.empty(base, dtype=torch.int32, device="cuda")
    st["val"] = torch.empty(base, dtype=torch.float32, device="cuda")
    return st


def run_influence(torch, st):
    """One build of the 3-field CSR: the fields' WET on the matrix's
    footprint (ray trace + smoothing) and their peak and count passes,
    overlapped on library streams with one host sync for all
    (`InfluenceMatrix.count_many`), then the fills into the shared CSR,
    overlapped on side streams (`InfluenceMatrix.fill_many`)."""
    from dosecuda_amd.engine import InfluenceMatrix
    InfluenceMatrix.count_many(st["infl"], st["d_rlsp"], st["wets"])
    InfluenceMatrix.fill_many(st["infl"], st["wets"], st["col"], st["val"])


def run_influence_serial(torch, st):
    """The same build one field at a time on one stream (per field: WET,
    peak and count with one host sync each; then the fills in turn): the
    kernels' own durations, which overlapped launches would share."""
    base = 0
    for f, w in zip(st["infl"], st["wets"]):
        f.wet(st["d_rlsp"], out=w)
        base += f.count(w, base)
    for f, w in zip(st["infl"], st["wets"]):
        f.fill(w, st["col"], st["val"])


def cone_steps(L0, spacing_mm):
    """Steps of the FP32 march `for (L = L0; L >= 0; L -= spacing/10)`."""
    sp = np.float32(spacing_mm) / np.float32(10.0)
    L, n = np.float32(L0), 0
    while L >= np.float32(0.0):
        n += 1
        L = np.float32(L - sp)
    return n


def timed(torch, fn, st, steps, warmup):
    for _ in range(warmup):
        fn(torch, st)
    torch.cuda.synchronize()
    from dosecuda_amd import engine
    engine.prof_reset()
    engine.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        fn(torch, st)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    engine.prof_enable(False)
    return (t1 - t0) * 1000.0 / steps


# ---------------------------------------------------------------------------
def cpu_baseline_cccs(st, budget_s):
    """FP64 oracle CCCS for beam 0 on a slab of z-planes through the field,
    scaled by active-voxel count to ms/beam."""
    from oracle import oracle as O
    pm, c = st["pm"], st["case"]
    dens = st["dens"]
    ga = st["angles"][0]
    g = O.geom(dens.shape, c.spacing, c.origin, c.iso, ga, 0.0, 0.0, np.float32(pm.primary_source_distance))
    t0 = time.perf_counter()
    raw, _ = O.raytrace(g, dens, float(np.float32(0.0012)))
    tr, el = O.terma(g, pm, st["mlc"], raw.astype(np.float32))
    t_pre = time.perf_counter() - t0
    t32, e32 = tr.astype(np.float32), el.astype(np.float32)
    active = ~(t32 <= 0.01)
    n_act = int(active.sum())
    nz = dens.shape[0]
    planes = 1
    mask = np.zeros(dens.shape, np.uint8)
    while True:
        mask[:] = 0
        k0 = nz // 2 - planes // 2
        mask[k0:k0 + planes] = 1
        t1 = time.perf_counter()
        O.ccc(g, pm, 100.0, t32, dens, e32, mask)
        dt = time.perf_counter() - t1
        n_s = int((active & (mask > 0)).sum())
        if dt > budget_s / 4 or planes >= nz:
            break
        planes = min(nz, planes * 2)
    per_vox = dt / max(n_s, 1)
    ms_beam = (t_pre + per_vox * n_act) * 1000.0
    return ms_beam, f"beam 0 of 7: ray trace + TERMA on the full 128^3 grid, CCCS on {planes} " \
                    f"z-planes ({n_s} of {n_act} active voxels), scaled by active voxels"


def _slab_mask(shape, planes):
    nz = shape[0]
    m = np.zeros(shape, np.uint8)
    k0 = max(0, nz // 2 - planes // 2)
    m[k0:k0 + planes] = 1
    return m


def cpu_baseline_cccs_1t(st, budget_s):
    """The CCCS baseline on one host thread: ray trace, TERMA and the cone
    superposition timed on a slab of z-planes through the field (TERMA of
    the other voxels, which the cones sample, from the multi-thread run) and
    scaled by voxel and active-voxel counts to ms/beam."""
    from oracle import oracle as O
    pm, c, dens = st["pm"], st["case"], st["dens"]
    ga = st["angles"][0]
    g = O.geom(dens.shape, c.spacing, c.origin, c.iso, ga, 0.0, 0.0, np.float32(pm.primary_source_distance))
    raw, _ = O.raytrace(g, dens, float(np.float32(0.0012)))
    tr, el = O.terma(g, pm, st["mlc"], raw.astype(np.float32))
    t32, e32 = tr.astype(np.float32), el.astype(np.float32)
    active = ~(t32 <= 0.01)
    n, n_act = dens.size, int(active.sum())
    nthr = O.get_threads()
    O.set_threads(1)
    try:
        planes = 1
        while True:
            mask = _slab_mask(dens.shape, planes)
            t0 = time.perf_counter()
            O.raytrace(g, dens, float(np.float32(0.0012)), mask=mask)
            t1 = time.perf_counter()
            O.terma(g, pm, st["mlc"], raw.astype(np.float32), mask=mask)
            t2 = time.perf_counter()
            O.ccc(g, pm, 100.0, t32, dens, e32, mask)
            t3 = time.perf_counter()
            if t3 - t0 > budget_s / 3 or planes >= dens.shape[0]:
                break
            planes = min(dens.shape[0], planes * 2)
    finally:
        O.set_threads(nthr)
    nv = int(mask.sum())
    na = int((active & (mask > 0)).sum())
    ms = ((t1 - t0 + t2 - t1) / nv * n + (t3 - t2) / max(na, 1) * n_act) * 1000.0
    return ms, f"one thread, beam 0 of 7: ray trace, TERMA and CCCS on {planes} z-planes ({nv} voxels, {na} " \
               f"active), scaled to the {dens.shape} grid ({n} voxels, {n_act} active)"


def cpu_baseline_impt_1t(st, budget_s):
    """The IMPT baseline on one host thread: ray trace on a slab of z-planes
    scaled by voxels, smoothing on the full grid, the pencil beam on a
    sample of spots scaled by voxel-spot pairs."""
    from oracle import oracle as O
    m, c = st["m"], st["case"]
    rlsp = st["rlsp"]
    D = (np.float32(m.VSADX) + np.float32(m.VSADY)) / np.float32(2.0)
    g = O.geom(rlsp.shape, c.spacing, c.origin, c.iso, 0.0, 0.0, 0.0, D)
    raw, _ = O.raytrace(g, rlsp, 0.0)
    wet32 = O.smooth(g, raw.astype(np.float32)).astype(np.float32)
    total_pairs = st["plan"].pairs(wet32)
    nthr = O.get_threads()
    O.set_threads(1)
    try:
        planes = 1
        while True:
            mask = _slab_mask(rlsp.shape, planes)
            t0 = time.perf_counter()
            O.raytrace(g, rlsp, 0.0, mask=mask)
            dt = time.perf_counter() - t0
            if dt > budget_s / 6 or planes >= rlsp.shape[0]:
                break
            planes = min(rlsp.shape[0], planes * 2)
        t_rt = dt / int(mask.sum()) * rlsp.size
        t0 = time.perf_counter()
        O.smooth(g, raw.astype(np.float32))
        t_sm = time.perf_counter() - t0
        spots = st["spots"]
        done, t_p = 0, 0.0
        for q in range(0, len(spots), 97):
            t1 = time.perf_counter()
            _, pairs = O.pencil(g, m.VSADX, m.VSADY, wet32, spots[q:q + 1], m.divergence_params, m.lut_depths,
                                m.lut_sigmas, m.lut_idds)
            t_p += time.perf_counter() - t1
            done += pairs
            if t_p > budget_s / 3:
                break
    finally:
        O.set_threads(nthr)
    t_field = t_rt + t_sm + t_p * total_pairs / max(done, 1)
    return total_pairs / t_field / 1e9, f"one thread: ray trace on {planes} z-planes scaled by voxels, smoothing " \
                                         f"on the full grid, pencil beam on {done} of {total_pairs} pairs"


def cpu_baseline_impt(st, budget_s):
    from oracle import oracle as O
    m, c = st["m"], st["case"]
    t0 = time.perf_counter()
    raw, sm, _ = O.proton_wet(st["rlsp"], c.spacing, c.origin, c.iso, 0.0, 0.0, m.VSADX, m.VSADY)
    t_wet = time.perf_counter() - t0
    wet32 = sm.astype(np.float32)
    # pencil beam on a subset of spots (one layer at a time) scaled by pairs
    spots = st["spots"]
    ids = np.unique(spots[:, 3])
    g = O.geom(wet32.shape, c.spacing, c.origin, c.iso, 0.0, 0.0, 0.0,
               (np.float32(m.VSADX) + np.float32(m.VSADY)) / np.float32(2.0))
    total_pairs = st["plan"].pairs(wet32)
    done_pairs, t_p = 0, 0.0
    for e in ids:
        sel = spots[spots[:, 3] == e]
        t1 = time.perf_counter()
        _, pairs = O.pencil(g, m.VSADX, m.VSADY, wet32, sel, m.divergence_params, m.lut_depths,
                            m.lut_sigmas, m.lut_idds)
        t_p += time.perf_counter() - t1
        done_pairs += pairs
        if t_p > budget_s:
            break
    t_field = t_wet + t_p * total_pairs / max(done_pairs, 1)
    return total_pairs / t_field / 1e9, f"WET on the full grid, pencil beam on {done_pairs} of " \
                                         f"{total_pairs} voxel-spot pairs, scaled by pairs"


def cpu_baseline_influence(st, budget_s):
    """FP64 oracle for one influence build: the WET of each field (ray trace
    + smoothing) on the full grid timed for field 0 and scaled by fields, and
    the single-spot pencil beam (every voxel of the grid, as a row needs) for
    a sample of spots, scaled by the spot count."""
    from oracle import oracle as O
    c, m = st["case"], st["m"]
    rlsp = st["d_rlsp"].cpu().numpy()
    g0 = st["fields"][0][0]
    t0 = time.perf_counter()
    _, sm, _ = O.proton_wet(rlsp, c.spacing, c.origin, c.iso, g0, 0.0, m.VSADX, m.VSADY)
    t_wet = time.perf_counter() - t0
    wet32 = sm.astype(np.float32)
    spots = st["fields"][0][1]
    n_all = sum(len(sp) for _, sp in st["fields"])
    done, t_sp = 0, 0.0
    for q in range(0, len(spots), max(1, len(spots) // 64)):
        one = np.array(spots[q], np.float32).reshape(1, 4)
        one[0, 2] = 1.0
        t1 = time.perf_counter()
        O.proton_dose(wet32, c.spacing, c.origin, c.iso, g0, 0.0, m.VSADX, m.VSADY, one, m.divergence_params,
                      m.lut_depths, m.lut_sigmas, m.lut_idds)
        t_sp += time.perf_counter() - t1
        done += 1
        if t_sp > budget_s:
            break
    ms = (t_wet * len(st["fields"]) + t_sp / max(done, 1) * n_all) * 1000.0
    return ms, f"WET of field 0 on the full grid (x {len(st['fields'])} fields), single-spot dose of " \
               f"{done} of {n_all} spots on the full grid, scaled by spots"


def _errs(got, ref):
    """(max, RMS) of |got - ref| relative to the oracle's Dmax."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    dm = float(np.abs(ref).max()) or 1.0
    d = np.abs(got - ref)
    return float(d.max()) / dm, float(np.sqrt(np.mean(d * d))) / dm


def oracle_errors_impt(st):
    """The field through the FP64 oracle: the raw WET end to end from the
    same RLSP; smoothing and pencil beam from the engine's float32 raw WET
    (the smoothing's neighbour set is a decision on the raw WET), against
    the engine's smoothed WET and dose."""
    import torch
    from dosecuda_amd import engine
    from oracle import oracle as O
    m, c = st["m"], st["case"]
    raw = torch.empty_like(st["d_rlsp"])
    sm_e = engine.proton_wet(st["d_rlsp"], c.spacing, c.origin, c.iso, 0.0, 0.0, m.VSADX, m.VSADY,
                             raw_out=raw)
    raw = raw.cpu().numpy()
    t0 = time.perf_counter()
    ref_raw, _, _ = O.proton_wet(st["rlsp"], c.spacing, c.origin, c.iso, 0.0, 0.0, m.VSADX, m.VSADY)
    g = O.geom(raw.shape, c.spacing, c.origin, c.iso, 0.0, 0.0, 0.0,
               (np.float32(m.VSADX) + np.float32(m.VSADY)) / np.float32(2.0))
    sm = O.smooth(g, raw)
    ref, _ = O.pencil(g, m.VSADX, m.VSADY, sm_e.cpu().numpy(), st["spots"], m.divergence_params, m.lut_depths,
                      m.lut_sigmas, m.lut_idds)
    mx, rms = _errs(st["dose"].cpu().numpy(), ref)
    wmx, wrms = _errs(raw, ref_raw)
    smx, srms = _errs(sm_e.cpu().numpy(), sm)
    return dict(max_rel=mx, rms_rel=rms, dmax=float(ref.max()), raw_wet_max_rel=wmx, raw_wet_rms_rel=wrms,
                smoothed_wet_max_rel=smx, smoothed_wet_rms_rel=srms, oracle_s=time.perf_counter() - t0,
                what="dose relative to the oracle's Dmax (pencil beam on the engine's smoothed WET); raw WET "
                     "end to end from the RLSP and smoothed WET from the engine's raw WET, relative to "
                     "their maxima")


def oracle_errors_cccs(torch, st):
    """Every beam of the plan through the FP64 oracle (ray trace, TERMA,
    72-cone superposition) from the same density, against the engine's
    per-beam dose, and the plan sum."""
    from dosecuda_amd import engine
    from oracle import oracle as O
    c, pm = st["case"], st["pm"]
    t0 = time.perf_counter()
    beams, psum, rsum = [], 0.0, 0.0
    terma = torch.empty_like(st["d_dens"])
    keep = np.ones(st["dens"].shape, bool)
    for ga in st["angles"]:
        got = engine.photon_dose(st["d_dens"], c.spacing, c.origin, c.iso, 100.0, ga, 0.0, 0.0, pm,
                                 st["mlc"], terma_out=terma).cpu().numpy()
        ref, stg = O.photon_dose(st["dens"], c.spacing, c.origin, c.iso, 100.0, ga, 0.0, 0.0, pm, st["mlc"],
                                 return_stages=True)
        # voxels whose TERMA sits on the 0.01 early-out to within rounding
        # (engine FP32 vs oracle FP64) are all-or-nothing: counted apart
        flip = (terma.cpu().numpy() <= np.float32(0.01)) != (stg["terma32"] <= np.float32(0.01))
        mx, rms = _errs(np.where(flip, 0.0, got), np.where(flip, 0.0, ref))
        beams.append(dict(gantry=float(ga), max_rel=mx, rms_rel=rms, early_out_flips=int(flip.sum()),
                          flip_terma_max_dev=float(np.abs(stg["terma"][flip] - 0.01).max()) if flip.any() else 0.0))
        keep &= ~flip
        psum = psum + got.astype(np.float64)
        rsum = rsum + ref
    mx, rms = _errs(np.where(keep, psum, 0.0), np.where(keep, rsum, 0.0))
    return dict(max_rel=max(b["max_rel"] for b in beams), rms_rel=max(b["rms_rel"] for b in beams),
                plan_max_rel=mx, plan_rms_rel=rms, early_out_flips=sum(b["early_out_flips"] for b in beams),
                per_beam=beams, oracle_s=time.perf_counter() - t0,
                what="each of the 7 beams end to end from the density, relative to that beam's oracle Dmax; "
                     "voxels whose TERMA lies on the 0.01 early-out to within rounding are counted in "
                     "early_out_flips and left out")


def oracle_errors_influence(torch, st, rows_per_field=8):
    """A sample of CSR rows against the oracle's single-spot (MU 1) dose on
    the same float32 WET: kept entries relative to the row peak, and the
    largest oracle value the truncation dropped."""
    from oracle import oracle as O
    c, m = st["case"], st["m"]
    col = st["col"].cpu().numpy().view(np.uint32)
    val = st["val"].cpu().numpy()
    n = st["d_rlsp"].numel()
    t0 = time.perf_counter()
    worst, rms_w, dropped, nrows, worst_row = 0.0, 0.0, 0.0, 0, None
    rng = np.random.default_rng(1003)
    for (g, spots), f, w in zip(st["fields"], st["infl"], st["wets"]):
        wet = w.cpu().numpy()
        row_spot, peak = f.info()
        rp = f.row_ptr
        for q in np.sort(rng.choice(len(row_spot), size=min(rows_per_field, len(row_spot)), replace=False)):
            one = np.array(spots[row_spot[q]], np.float32).reshape(1, 4)
            one[0, 2] = 1.0
            ref, _ = O.proton_dose(wet, c.spacing, c.origin, c.iso, g, 0.0, m.VSADX, m.VSADY, one,
                                   m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
            ref = ref.reshape(-1)
            pk = float(ref.max()) or 1.0
            got = np.zeros(n)
            a, b = int(rp[q]), int(rp[q + 1])
            got[col[a:b]] = val[a:b]
            kept = got != 0
            d = np.abs(got[kept] - ref[kept]) / pk
            worst = max(worst, float(d.max()) if d.size else 0.0)
            rms_w = max(rms_w, float(np.sqrt(np.mean(d * d))) if d.size else 0.0)
            dr = float(ref[~kept].max()) / pk if (~kept).any() else 0.0
            if dr > dropped:
                dropped = dr
                worst_row = dict(gantry=float(g), row=int(q), spot=[float(v) for v in one[0]], entries=b - a,
                                 engine_peak=float(peak[q]), oracle_peak=pk,
                                 oracle_peak_voxel=int(np.argmax(ref)))
            nrows += 1
    return dict(rows=nrows, max_rel=worst, rms_rel=rms_w, max_dropped_rel=dropped, worst_dropped_row=worst_row,
                oracle_s=time.perf_counter() - t0,
                what="sampled rows on the engine's float32 WET, relative to each row's oracle peak; "
                     "max_dropped_rel is the largest oracle value left out by the 1e-4 truncation")


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU per process; a run over several "
                                   "GPUs is not implemented (--gpus must be 1)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch
    from dosecuda_amd import _lib, engine
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    _lib.load()
    if a.no_tune:
        from dosecuda_amd import autotune
        log(f"tuning cache {autotune.cache_path()}: {autotune.load()} choices loaded")
    res = {}
    cpu = {}
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if not a.no_cpu:
        from oracle import oracle as O
        threads = O.get_threads()          # the oracle's OpenMP threads (the cores it uses)

    if a.workload in ("all", "impt"):
        log("setting up IMPT 100^3 field")
        st = setup_impt(torch)
        if not a.no_tune:
            from dosecuda_amd import autotune
            c, m = st["case"], st["m"]
            tw = autotune.tune_proton_wet(st["d_rlsp"], c.spacing, c.origin, c.iso, 0.0, 0.0, m, reps=5)
            run_impt(torch, st)
            tp = autotune.tune_pencil(st["plan"], st["wet"], reps=5)
            res["impt_tuning"] = dict(raytrace=tw["raytrace"], smooth=tw["smooth"], pencil_tile=tp["tile"],
                                      pencil_lut_in_lds=tp["lut_in_lds"])
            log(f"IMPT tuning: {res['impt_tuning']}")
        ms = timed(torch, run_impt, st, a.steps, a.warmup)
        wet = st["wet"].cpu().numpy()
        pairs = st["plan"].pairs(wet)
        kt = {k: engine.prof_get(k) for k in ("pad", "raytrace", "smooth", "pencil")}
        pen_ms = kt["pencil"][0] / max(kt["pencil"][1], 1)
        # k_pencil is VALU-bound: 11 full-rate VALU ops + 2 quarter-rate
        # v_exp_f32 = 19 lane-op slots per pair; 256 CUs x 64 lane-ops / clock
        valu_peak = 256 * 64 * 2.4e9 / 19 / 1e9
        res["impt"] = dict(ms_per_field=ms, pairs=pairs, gvoxel_spot_per_s=pairs / (ms * 1e-3) / 1e9,
                           pencil_ms=pen_ms, pencil_gvoxel_spot_per_s=pairs / (pen_ms * 1e-3) / 1e9,
                           pencil_valu_roofline=dict(bound="VALU", achieved=pairs / (pen_ms * 1e-3) / 1e9,
                                                     peak=valu_peak, unit="Gpair/s",
                                                     frac=pairs / (pen_ms * 1e-3) / 1e9 / valu_peak),
                           kernels_ms={k: v[0] / max(v[1], 1) for k, v in kt.items()})
        log(f"IMPT: {ms:.3f} ms/field, {res['impt']['gvoxel_spot_per_s']:.1f} Gvoxel-spot/s "
            f"(pencil {pen_ms:.3f} ms)")
        if not a.no_cpu:
            v, sample = cpu_baseline_impt(st, a.cpu_seconds / 2)
            cpu["impt"] = dict(value=v, unit="Gvoxel-spot/s", sample=sample)
            log(f"IMPT CPU oracle: {v:.4f} Gvoxel-spot/s")
            v1, sample1 = cpu_baseline_impt_1t(st, a.cpu_seconds)
            cpu["impt"]["single_thread"] = dict(value=v1, unit="Gvoxel-spot/s", cores=1, sample=sample1)
            log(f"IMPT CPU oracle, one thread: {v1:.5f} Gvoxel-spot/s")
            res["impt"]["oracle_error"] = oracle_errors_impt(st)
            log(f"IMPT vs FP64 oracle: {res['impt']['oracle_error']}")
        st["plan"].close()
        del st

    if a.workload in ("all", "cccs"):
        log("setting up CCCS 128^3 x 7 beams")
        st = setup_cccs(torch)
        if not a.no_tune:
            from dosecuda_amd import autotune
            c = st["case"]
            # the choices are keyed by shape and dominant beam axis: one
            # gantry angle per axis the 7 beams use
            for g in (0.0, 51.42857142857143):
                tw = autotune.tune_photon(st["d_dens"], c.spacing, c.origin, c.iso, 100.0, g, 0.0, 0.0, st["pm"],
                                          st["mlc"], reps=5)
                res.setdefault("cccs_tuning", {})[f"gantry {g:.1f}"] = dict(
                    raytrace=tw["raytrace"], ccc_layout=tw["ccc_layout"],
                    ccc_ms={"pairs": tw["ccc_ms"][0], "quads": tw["ccc_ms"][1]})
            log(f"CCCS tuning: {res['cccs_tuning']}")
        ms = timed(torch, run_cccs, st, a.steps, a.warmup)
        nb = len(st["angles"])
        kt = {k: engine.prof_get(k) for k in ("pad", "raytrace", "terma", "pack_tr", "ccc_compact", "ccc",
                                              "ccc_finish")}
        ccc_ms = kt["ccc"][0] / max(kt["ccc"][1], 1)
        # algorithmic bytes of one k_ccc launch: the padded {TERMA, density}
        # volume once (8 B per padded voxel) + per active voxel: its list
        # index and its 12 azimuth sums (52 B; k_ccc_finish adds them,
        # the electron term and writes the dose)
        c = st["case"]
        from dosecuda_amd import _lib as L
        n = st["dens"].size
        npad = (st["dens"].shape[0] + 3) * (st["dens"].shape[1] + 3) * (st["dens"].shape[2] + 3)
        # active voxels of the last beam
        terma = torch.empty_like(st["d_dens"])
        engine.photon_dose(st["d_dens"], c.spacing, c.origin, c.iso, 100.0, st["angles"][-1], 0.0,
                           0.0, st["pm"], st["mlc"], out=st["out"], terma_out=terma)
        n_act = int((~(terma <= 0.01)).sum().item())
        bytes_launch = 8 * npad + 52 * n_act
        achieved = bytes_launch / (ccc_ms * 1e-3) / 1e9
        samples = n_act * 12 * sum(cone_steps(L0, c.spacing) for L0 in st["pm"].kernel[:, 5])
        # k_ccc is bound by its trilinear gathers: 8 corners x {TERMA,
        # density} = 64 bytes per cone sample through the vector L1, whose
        # return path moves 64 B per clock per CU
        cus, clk = 256, 2.4e9
        l1_peak = cus * 64 * clk / 1e12
        l1_ach = samples * 64 / (ccc_ms * 1e-3) / 1e12
        res["cccs"] = dict(ms_per_step=ms, ms_per_beam=ms / nb, ccc_ms=ccc_ms, n_active=n_act,
                           cone_samples_per_beam=samples,
                           gsamples_per_s=samples / (ccc_ms * 1e-3) / 1e9,
                           gather_roofline=dict(bound="L1 gather", bytes_per_sample=64, achieved=l1_ach,
                                                peak=l1_peak, unit="TB/s", frac=l1_ach / l1_peak),
                           kernels_ms={k: v[0] / max(v[1], 1) for k, v in kt.items()},
                           roofline=dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS,
                                         unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                                         traffic=CCC_TRAFFIC_BYTES,
                                         traffic_source="profiled constant: rocprofv3 --pmc FETCH_SIZE / "
                                                        "WRITE_SIZE of k_ccc in this bench, calibrated "
                                                        "(profiles/hbm_traffic_cccs128_7beams.md)"))
        log(f"CCCS: {ms / nb:.3f} ms/beam (ccc kernel {ccc_ms:.3f} ms, {n_act} active voxels)")
        if not a.no_cpu:
            v, sample = cpu_baseline_cccs(st, a.cpu_seconds)
            cpu["cccs"] = dict(value=v, unit="ms/beam", sample=sample)
            log(f"CCCS CPU oracle: {v:.1f} ms/beam")
            v1, sample1 = cpu_baseline_cccs_1t(st, a.cpu_seconds)
            cpu["cccs"]["single_thread"] = dict(value=v1, unit="ms/beam", cores=1, sample=sample1)
            log(f"CCCS CPU oracle, one thread: {v1:.0f} ms/beam")
            res["cccs"]["oracle_error"] = oracle_errors_cccs(torch, st)
            log(f"CCCS vs FP64 oracle: max {res['cccs']['oracle_error']['max_rel']:.2e} "
                f"rms {res['cccs']['oracle_error']['rms_rel']:.2e}")

    if a.workload in ("all", "influence"):
        log("setting up the batched influence matrix: (160,256,256) @ 2 mm, 3 fields, ~15k spots")
        st = setup_influence(torch, tune=not a.no_tune)
        if st["tuning"]:
            log(f"influence tuning (spots per workgroup, row chunks): {st['tuning']}")
        ms = timed(torch, run_influence, st, a.influence_steps, 1)
        # kernel durations (and the fill's roofline) from the one-field-at-a-
        # time build: the overlapped launches' event intervals share the GPU
        ms_serial = timed(torch, run_influence_serial, st, a.influence_steps, 1)
        kt = {k: engine.prof_get(k) for k in ("raytrace", "smooth", "influence_rows", "influence_bounds",
                                              "influence_axis", "influence_peak", "influence_count",
                                              "influence_fill")}
        fill_ms = kt["influence_fill"][0] / max(kt["influence_fill"][1], 1)
        nvox = st["d_rlsp"].numel()
        n_rows = sum(len(sp) for _, sp in st["fields"])
        # algorithmic bytes of one fill launch (one field on average): the
        # field's entries (u32 column + f32 value) and one read of its WET
        bytes_launch = (8 * st["nnz"] + 4 * nvox * len(st["fields"])) / len(st["fields"])
        ach = bytes_launch / (fill_ms * 1e-3) / 1e9
        # algorithmic bytes of a whole build: the CSR (u32 column + f32 value
        # per entry, i64 row pointers) and per field the RLSP read plus the
        # WET written and read once
        build_bytes = 8 * st["nnz"] + 8 * (n_rows + 1) + len(st["fields"]) * 3 * 4 * nvox
        build_ach = build_bytes / (ms * 1e-3) / 1e9
        res["influence"] = dict(ms_per_build=ms, ms_per_build_one_field_at_a_time=ms_serial, nnz=st["nnz"],
                                rows=n_rows, tuning=st["tuning"],
                                build_roofline=dict(bound="hbm", bytes=build_bytes, achieved=build_ach,
                                                    peak=HBM_PEAK_GBS, unit="GB/s", frac=build_ach / HBM_PEAK_GBS),
                                fill_ms=fill_ms, csr_bytes=8 * st["nnz"] + 8 * (n_rows + 1),
                                kernels_ms={k: v[0] / max(v[1], 1) for k, v in kt.items()},
                                roofline=dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS,
                                              unit="GB/s", frac=ach / HBM_PEAK_GBS,
                                              traffic=INFL_FILL_TRAFFIC_BYTES,
                                              traffic_source="profiled constant: rocprofv3 --pmc FETCH_SIZE / "
                                                             "WRITE_SIZE of the fill, calibrated by a 256 MiB copy "
                                                             "(profiles/kernel_stats_influence_3fields.md)"))
        sweep = []
        for f in st["infl"]:
            c, rad = f.stats()
            sweep.append(dict(c, mean_radius_mm=float(rad.mean())))
        res["influence"]["sweep"] = sweep
        fps = []
        for f, w in zip(st["infl"], st["wets"]):
            box, on, fb = f.footprint()
            fps.append(dict(box_mm=[float(v) for v in box], in_use=on, fallbacks=fb,
                            voxel_fraction=float((w < 1e29).float().mean())))
        res["influence"]["footprint"] = fps
        log(f"influence footprints: {fps}")
        if not a.no_cpu:
            res["influence"]["oracle_error"] = oracle_errors_influence(torch, st)
            log(f"influence rows vs FP64 oracle: {res['influence']['oracle_error']}")
            v, sample = cpu_baseline_influence(st, a.cpu_seconds / 2)
            cpu["influence"] = dict(value=v, unit="ms/build", sample=sample)
            log(f"influence CPU oracle: {v:.0f} ms/build")
        log(f"influence sweep counters: {sweep}")
        log(f"influence: {ms:.2f} ms/build ({ms_serial:.2f} one field at a time), nnz {st['nnz']}, "
            f"fill {fill_ms:.3f} ms/field "
            f"({ach:.0f} GB/s = {100 * ach / HBM_PEAK_GBS:.1f}% of HBM)")
        for f in st["infl"]:
            f.close()
        del st

    if a.workload in ("all", "surrogate"):
        from dosecuda_amd import autotune, surrogate as S
        from dosecuda_amd.beam_models import PhotonBeamModel
        if not os.path.exists(S.WEIGHTS):
            log(f"surrogate: no trained weights at {S.WEIGHTS}; skipped")
        else:
            net = S.load(S.WEIGHTS)
            if not a.no_tune:
                t0 = time.time()
                # the plan's 7 beams run through the network as one batch
                autotune.tune_surrogate(net, (64, 64, 64), reps=5, batch=7)
                log(f"surrogate: convolution variants tuned ({time.time() - t0:.1f} s)")
            log("surrogate: 7-beam plan, checked surrogate vs plain CCCS")
            r = S.evaluate(net, PhotonBeamModel(), stride=S.DEFAULT_CHECK[0], tol=S.DEFAULT_CHECK[1],
                           dilate=S.DEFAULT_CHECK[2], reps=5)
            conv_ms_single = float(np.mean([b["surrogate_stage_ms"]["nn_conv3d"] for b in r["beams"]]))
            conv_ms = r["batch_conv_ms_per_beam"]
            tf = r["net_gflop"] / (conv_ms * 1e-3) / 1e3
            # the whole surrogate path per beam from the density (TERMA with one
            # ray sample per voxel, checked surrogate batched over the plan)
            sur_beam = r["path_ms_per_beam"]
            res["surrogate"] = dict(
                speedup_vs_plain_beam=r["plain_beam_ms"] / sur_beam,
                surrogate_beam_ms=sur_beam, plain_beam_ms=r["plain_beam_ms"],
                speedup_target=10.0,
                min_beam_gamma_path=r["path_min_gamma"], plan_gamma_path=r["path_plan_gamma"],
                superposition_only_ms_per_beam=r["graph_checked_ms_per_beam"],
                weights=os.path.relpath(S.WEIGHTS, ROOT), params=r["params"], stride=r["stride"], tol=r["tol"],
                dilate=r["dilate"],
                plan_gamma_checked=r["plan_gamma_checked"], min_beam_gamma_checked=r["min_gamma_checked"],
                plan_gamma_net_only=r["plan_gamma_net"], min_beam_gamma_net_only=r["min_gamma_net"],
                max_err_checked=max(b["max_err_checked"] for b in r["beams"]),
                rms_err_checked=max(b["rms_err_checked"] for b in r["beams"]),
                mean_fallback_fraction=r["mean_fallback_fraction"],
                checked_ms_per_beam=r["graph_checked_ms_per_beam"],
                checked_ms_per_beam_unbatched=r["graph_single_ms_per_beam"],
                batch_equal_single=r["batch_equal_single"],
                plain_superposition_ms_per_beam=r["plain_superposition_ms"],
                speedup_vs_plain_superposition=r["speedup_superposition"],
                mfma=dict(net_gflop=r["net_gflop"], conv_ms_per_beam=conv_ms, achieved_tflops=tf, peak_tflops=2500.0,
                          frac=tf / 2500.0, conv_ms_per_beam_unbatched=conv_ms_single),
                graph_replay_equal=r["graph_replay_equal"])
            log(f"surrogate: whole beam {sur_beam:.3f} ms vs plain {r['plain_beam_ms']:.3f} ms "
                f"({r['plain_beam_ms'] / sur_beam:.2f}x, min beam gamma {r['path_min_gamma']:.5f}); checked superposition "
                f"{r['graph_checked_ms_per_beam']:.3f} ms/beam vs {r['plain_superposition_ms']:.3f} "
                f"plain superposition ({r['speedup_superposition']:.2f}x), plan gamma {r['plan_gamma_checked']:.5f}, "
                f"min beam {r['min_gamma_checked']:.5f}, fallback {r['mean_fallback_fraction']:.3f}, "
                f"MFMA {tf:.0f} TFLOP/s")

    if a.workload == "surrogate_path":
        from dosecuda_amd import autotune, surrogate as S
        from dosecuda_amd.beam_models import PhotonBeamModel
        net = S.load(S.WEIGHTS)
        if not a.no_tune:
            autotune.tune_surrogate(net, (64, 64, 64), reps=5, batch=7)
        r = S.time_plan_path(net, PhotonBeamModel(), reps=a.steps)
        res["surrogate_path"] = r
        log(f"surrogate path: {r['ms_per_beam']:.4f} ms/beam over {a.steps} replays, min beam gamma "
            f"{r['min_gamma']:.5f}")

    if a.workload in ("all", "mixed"):
        from dosecuda_amd import mixed, phantoms, surrogate as S
        from dosecuda_amd.beam_models import PhotonBeamModel, ProtonBeamModel
        pm, xm = ProtonBeamModel(), PhotonBeamModel()
        cases = phantoms.mixed_cases(pm, xm, seed=1005, n_cases=a.mixed_cases)
        nat = S.NativeNet(S.load(S.WEIGHTS)) if os.path.exists(S.WEIGHTS) else None
        log(f"mixed: {len(cases)} cases, surrogate {'on' if nat else 'off (no weights)'}")
        st, tol, dil = S.DEFAULT_CHECK
        W = mixed.MixedWorkload(cases, pm, xm, surrogate=nat, streams=4, stride=st, tol=tol, dilate=dil, log=log)
        cnt = W.counts()
        log(f"mixed: set up in {W.setup_s:.1f} s: {cnt}")
        t0 = time.time()
        W.run(graph=True)          # eager warm-up (sizes the workspaces, times the jobs), capture, launch
        torch.cuda.synchronize()
        log(f"mixed: warm-up + capture + first replay {time.time() - t0:.1f} s")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(a.mixed_steps):
            W.run(graph=True)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / a.mixed_steps
        errs = W.errors()
        imrt = [e for e in errs if e["kind"] == "imrt"]
        impt = [e for e in errs if e["kind"] == "impt"]
        # the IMPT fields' peak/count passes run once at set-up (they size the
        # CSR arrays the graph fills); both rates are reported
        count_ms = W.count_s * 1e3
        res["mixed"] = dict(
            counts=cnt, ms_per_batch=ms, cases_per_s=cnt["cases"] / (ms * 1e-3),
            count_passes_ms=count_ms, ms_per_batch_with_counts=ms + count_ms,
            cases_per_s_with_counts=cnt["cases"] / ((ms + count_ms) * 1e-3),
            nnz_per_s=cnt["nnz"] / (ms * 1e-3), voxel_cps_per_s=cnt["voxel_cps"] / (ms * 1e-3),
            surrogate=nat is not None,
            imrt_min_gamma=min((e["gamma"] for e in imrt), default=None),
            imrt_max_err=max((e["max_err"] for e in imrt), default=None),
            impt_max_err_csr_vs_dense=max((e["max_err"] for e in impt), default=None),
            per_case=errs)
        log(f"mixed: {ms / 1e3:.2f} s per batch ({(ms + count_ms) / 1e3:.2f} s with the IMPT count passes, "
            f"{res['mixed']['cases_per_s_with_counts']:.1f} cases/s), {res['mixed']['cases_per_s']:.1f} cases/s, min IMRT gamma "
            f"{res['mixed']['imrt_min_gamma']}, max IMPT CSR error {res['mixed']['impt_max_err_csr_vs_dense']}")
        W.close()

    if not a.no_tune:
        from dosecuda_amd import autotune
        if autotune.entries():
            log(f"tuning cache saved to {autotune.save()}")
    cc = res.get("cccs")
    line = {
        "metric": "IMPT Gvoxel-spot/s + CCCS ms/beam; surrogate vs plain at 2%/2mm gamma>=99%",
        "value": cc["ms_per_beam"] if cc else None,
        "unit": "ms/beam",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": cc["ms_per_step"] if cc else None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded phantoms; unchanged DoseCUDA beam-model tables)",
        "config": {"workload": "plain IMRT CCCS, 128^3 @ 2.5 mm water + lung ellipsoid + bone "
                               "cylinder (seed 1002), 7 coplanar 6 MV beams, 10x10 cm MLC field, "
                               "VarianTrueBeamHF",
                   "beams": 7, "grid": [128, 128, 128], "spacing_mm": 2.5},
        "roofline": cc["roofline"] if cc else None,
        "cpu_baseline": ({"value": cpu["cccs"]["value"], "unit": "ms/beam", "cores": threads,
                          "kind": "port", "sample": cpu["cccs"]["sample"],
                          "single_thread": cpu["cccs"].get("single_thread")} if "cccs" in cpu else None),
        "oracle_error": {k: ({kk: vv for kk, vv in res[k]["oracle_error"].items() if kk != "per_beam"}
                             if res.get(k) and "oracle_error" in res[k] else None)
                         for k in ("impt", "cccs", "influence")},
        "impt": res.get("impt"),
        "tuning": {k: res[k] for k in ("impt_tuning", "cccs_tuning") if k in res},
        "impt_cpu_baseline": ({**cpu["impt"], "cores": threads, "kind": "port"} if "impt" in cpu else None),
        "influence_cpu_baseline": ({**cpu["influence"], "cores": threads, "kind": "port"}
                                   if "influence" in cpu else None),
        "cccs_detail": cc,
        "influence": res.get("influence"),
        "surrogate": res.get("surrogate"),
        "surrogate_path": res.get("surrogate_path"),
        "mixed": res.get("mixed"),
        "autotuned": not a.no_tune,
    }
    print(json.dumps(line))
    if a.copy_calib_mb:
        n = a.copy_calib_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {a.copy_calib_mb} MiB read + {a.copy_calib_mb} MiB written")


if __name__ == "__main__":
    main()
