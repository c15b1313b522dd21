"""GPU: the learned-surrogate and mixed workloads at their full sizes.

* Surrogate plan: the seeded 128^3 lung/bone phantom at 2.5 mm (seed 1002)
  with 7 coplanar 10x10 cm beams.  Each beam's checked surrogate dose
  (network, lattice check at stride 3 with 1% of Dmax tolerance and face
  dilation, plain fallback) passes 2%/2 mm global gamma >= 99% against the
  plain CCCS of the same beam; the beams batched through the network equal
  the single-beam doses bit for bit.  The fallback fraction and the
  network-alone gamma are printed.
* Mixed batch: the 8 largest of the 64 seeded cases (seed 1005, sides
  96-256 voxels at 2-3 mm; 4 IMPT cases as CSR influence matrices and 4 IMRT
  cases as checked-surrogate control points) run eagerly and then as one HIP
  graph replayed twice, equal bit for bit.  One full-size IMRT control point
  (plain CCCS) and one full-size IMPT field (raw WET end to end, smoothing of
  the engine's raw WET, then the dense dose and sampled CSR rows of its
  first energy layer) are checked against the FP64 oracle within 1e-4 of
  Dmax (rows: of each row's peak; WET: 1e-5 of its maximum).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu
F = np.float32
TOL = 1e-4


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def test_surrogate_plan_128_7beams(photon_model):
    t = _torch()
    from dosecuda_amd import engine, phantoms, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained surrogate weights")
    case = phantoms.imrt_lung_bone_case(seed=1002, n=128, spacing=2.5)
    rho = t.from_numpy(phantoms.to_density(case)).cuda()
    mlc = square_field_mlc(photon_model, 10.0, 10.0)
    sp, org, iso = case.spacing, case.origin, case.iso
    sur = S.SurrogateCCCS(S.load(S.WEIGHTS), photon_model, stride=3, tol=0.01, dilate=1)
    angles = phantoms.imrt_beams(7)
    termas, els, plains, singles = [], [], [], []
    for ga in angles:
        print(f"[case] surrogate 128^3 beam gantry={ga:.2f}", flush=True)
        terma, el = t.empty_like(rho), t.empty_like(rho)
        plain = engine.photon_dose(rho, sp, org, iso, 100.0, ga, 0.0, 0.0, photon_model, mlc,
                                   terma_out=terma, electron_out=el)
        tm = {}
        dc = sur.dose(terma, rho, el, sp, org, iso, 100.0, ga, 0.0, 0.0, timings=tm)
        dn = sur.predict(terma, rho, el, ga, 0.0, 100.0, sp)
        g_chk, _ = S.gamma(dc, plain, sp)
        g_net, _ = S.gamma(dn, plain, sp)
        frac = tm["fallback_voxels"] / max(tm["active_voxels"], 1)
        print(f"[parity] surrogate beam {ga:.2f}: gamma checked {g_chk:.5f} network alone {g_net:.5f} "
              f"fallback {frac:.3f} of active voxels", flush=True)
        assert g_chk >= 0.99, ga
        termas.append(terma)
        els.append(el)
        plains.append(plain)
        singles.append(dc.clone())
    outs = sur.dose_batch(termas, rho, els, sp, org, [(iso, 100.0, ga, 0.0, 0.0) for ga in angles])
    t.cuda.synchronize()
    for o, s in zip(outs, singles):
        assert t.equal(o, s)
    g_plan, _ = S.gamma(sum(outs), sum(plains), sp)
    print(f"[parity] surrogate plan gamma {g_plan:.5f}", flush=True)
    assert g_plan >= 0.99
    # the whole surrogate path from the density (TERMA with one ray sample
    # per voxel spacing, the default check), every beam against its plain dose
    st, tol, dil = S.DEFAULT_CHECK
    sur2 = S.SurrogateCCCS(S.load(S.WEIGHTS), photon_model, stride=st, tol=tol, dilate=dil)
    cps = [(iso, 100.0, ga, 0.0, 0.0) for ga in angles]
    full = sur2.plan_doses(rho, sp, org, cps, [mlc for _ in angles])
    for ga, o, p in zip(angles, full, plains):
        g_path, _ = S.gamma(o, p, sp)
        print(f"[parity] surrogate whole path beam {ga:.2f}: gamma {g_path:.5f}", flush=True)
        assert g_path >= 0.99, ga


def _largest(cases, kind, n):
    sel = [(int(np.prod(c.shape)), i) for i, c in enumerate(cases) if c.kind == kind]
    return [i for _, i in sorted(sel, reverse=True)[:n]]


def test_mixed_largest_cases_graph_and_oracle(proton_model, photon_model):
    t = _torch()
    from dosecuda_amd import engine, mixed, phantoms, surrogate as S
    pm, xm = proton_model, photon_model
    cases = phantoms.mixed_cases(pm, xm, seed=1005, n_cases=64)
    pick = sorted(_largest(cases, "impt", 4) + _largest(cases, "imrt", 4))
    sub = [cases[i] for i in pick]
    print(f"[case] mixed largest: {[(i, c.kind, c.shape) for i, c in zip(pick, sub)]}", flush=True)
    nat = S.NativeNet(S.load(S.WEIGHTS)) if os.path.exists(S.WEIGHTS) else None
    st, tol, dil = S.DEFAULT_CHECK
    W = mixed.MixedWorkload(sub, pm, xm, surrogate=nat, streams=4, stride=st, tol=tol, dilate=dil)
    W.run(graph=False)
    t.cuda.synchronize()
    eager = [(it["dose"].clone() if it["kind"] == "imrt" else [f["val"][:f["nnz"]].clone() for f in it["fields"]])
             for it in W.items]
    for rep in range(2):
        W.run(graph=True)
        t.cuda.synchronize()
        for e, it in zip(eager, W.items):
            if it["kind"] == "imrt":
                assert t.equal(e, it["dose"]), rep
            else:
                assert all(t.equal(a, f["val"][:f["nnz"]]) for a, f in zip(e, it["fields"])), rep
    print(f"[parity] mixed largest: graph replays equal the eager batch", flush=True)

    # one full-size IMRT control point, plain CCCS, against the oracle
    imrt = [it for it in W.items if it["kind"] == "imrt"]
    it = min(imrt, key=lambda x: int(np.prod(x["case"].shape)))
    c = it["case"]
    g, mu, mlc = c.cps[0]
    rho = it["rho"]
    got = engine.photon_dose(rho, c.spacing, np.zeros(3), it["iso"], mu, g, 0.0, 0.0, xm, mlc).cpu().numpy()
    ref = O.photon_dose(rho.cpu().numpy(), c.spacing, np.zeros(3), it["iso"], mu, g, 0.0, 0.0, xm, mlc)
    mx, rms = rel_err(got, ref)
    print(f"[parity] mixed IMRT cp shape={c.shape} vs oracle: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= TOL

    # one full-size IMPT field: WET, dense dose of its first layer, CSR rows of that layer
    impt = [it for it in W.items if it["kind"] == "impt"]
    it = min(impt, key=lambda x: int(np.prod(x["case"].shape)))
    c, fl = it["case"], it["fields"][0]
    # raw WET end to end; the smoothing's neighbour set is decided from the
    # raw WET (IMPTClasses.cu:160-186), so the oracle smooths the engine's
    # float32 raw WET, as in the 100^3 field test
    raw = t.empty_like(it["rlsp"])
    sm_e = engine.proton_wet(it["rlsp"], c.spacing, np.zeros(3), it["iso"], fl["gantry"], 0.0, pm.VSADX,
                             pm.VSADY, raw_out=raw)
    assert t.equal(sm_e, fl["wet"])
    ref_raw, _, _ = O.proton_wet(it["rlsp"].cpu().numpy(), c.spacing, np.zeros(3), it["iso"], fl["gantry"],
                                 0.0, pm.VSADX, pm.VSADY)
    raw = raw.cpu().numpy()
    mx, rms = rel_err(raw, ref_raw)
    print(f"[parity] mixed IMPT raw WET shape={c.shape} vs oracle: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= 1e-5
    g = O.geom(c.shape, c.spacing, np.zeros(3), it["iso"], fl["gantry"], 0.0, 0.0,
               (F(pm.VSADX) + F(pm.VSADY)) / F(2.0))
    wet = fl["wet"].cpu().numpy()
    mx, rms = rel_err(wet, O.smooth(g, raw))
    print(f"[parity] mixed IMPT smoothed WET vs oracle: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= 1e-5
    spots = fl["spots"]
    layer = spots[spots[:, 3] == spots[0, 3]]
    plan = engine.ProtonPlan(0, c.shape, c.spacing, np.zeros(3), it["iso"], fl["gantry"], 0.0, pm, layer)
    got = plan.dose(fl["wet"]).cpu().numpy()
    plan.close()
    ref, _ = O.proton_dose(wet, c.spacing, np.zeros(3), it["iso"], fl["gantry"], 0.0, pm.VSADX, pm.VSADY, layer,
                           pm.divergence_params, pm.lut_depths, pm.lut_sigmas, pm.lut_idds)
    mx, rms = rel_err(got, ref)
    print(f"[parity] mixed IMPT layer of {len(layer)} spots vs oracle: max {mx:.3e} rms {rms:.3e}", flush=True)
    assert mx <= TOL
    row_spot, _ = fl["f"].info()
    rp = np.asarray(fl["f"].row_ptr, np.int64)
    col = fl["col"].cpu().numpy().view(np.uint32)
    val = fl["val"].cpu().numpy()
    n = int(np.prod(c.shape))
    in_layer = [q for q in range(len(row_spot)) if spots[row_spot[q], 3] == spots[0, 3]]
    for q in in_layer[:: max(1, len(in_layer) // 4)][:4]:
        one = np.array(spots[row_spot[q]], F).reshape(1, 4)
        one[0, 2] = 1.0
        r1, _ = O.proton_dose(wet, c.spacing, np.zeros(3), it["iso"], fl["gantry"], 0.0, pm.VSADX, pm.VSADY, one,
                              pm.divergence_params, pm.lut_depths, pm.lut_sigmas, pm.lut_idds)
        r1 = r1.reshape(-1)
        pk = float(r1.max())
        row = np.zeros(n)
        a, b = int(rp[q] - rp[0]), int(rp[q + 1] - rp[0])
        row[col[a:b]] = val[a:b]
        kept = row != 0
        err = float(np.abs(row[kept] - r1[kept]).max()) / pk
        dropped = float(r1[~kept].max()) / pk
        print(f"[parity] mixed CSR row {q}: {b - a} entries, max {err:.3e} of peak, largest dropped "
              f"{dropped:.3e}", flush=True)
        assert err <= TOL and dropped <= 1.1e-4
    W.close()
