"""GPU: the batching scheduler reproduces the single-call paths bit for bit
(photon plain and surrogate control points summed per case in job order,
proton WET + dense dose, influence fills), eagerly and as a replayed HIP
graph across several streams."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _photon_cases(t, rng, n_cases=5):
    out = []
    for c in range(n_cases):
        shape = tuple(int(v) for v in rng.integers(24, 56, 3))
        s = float(rng.uniform(2.0, 3.0))
        rho = t.full(shape, 1.00653, device="cuda")
        z0, y0, x0 = (int(v) for v in rng.integers(2, 12, 3))
        rho[z0:z0 + 8, y0:y0 + 10, x0:x0 + 9] = float(rng.uniform(0.25, 0.4))
        iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
        beams = [(float(rng.uniform(0, 360)), float(rng.uniform(3, 8))) for _ in range(int(rng.integers(1, 4)))]
        out.append((rho, s, iso, beams))
    return out


def test_scheduler_photon_eager_and_graph(photon_model):
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import square_field_mlc
    from dosecuda_amd.scheduler import Scheduler
    pm = photon_model
    rng = np.random.default_rng(11)
    cases = _photon_cases(t, rng)
    print(f"[case] scheduler photon, {len(cases)} cases", flush=True)
    sched = Scheduler(0, streams=3)
    doses, refs = [], []
    for rho, s, iso, beams in cases:
        d = t.full_like(rho, -1.0)
        cid = sched.add_case(rho, s, np.zeros(3), d)
        ref = t.zeros_like(rho)
        for ga, fs in beams:
            mlc = square_field_mlc(pm, fs, fs)
            sched.add_photon(cid, iso, 100.0, ga, 0.0, 0.0, pm, mlc)
            ref += engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, ga, 0.0, 0.0, pm, mlc)
        doses.append(d)
        refs.append(ref)
    sched.run(graph=False)
    t.cuda.synchronize()
    for d, r in zip(doses, refs):
        assert t.equal(d, r)
    info = sched.info()
    assert info["cases"] == len(cases) and all(ms > 0 for ms in info["job_ms"])
    assert len(set(info["case_stream"])) > 1
    for rep in range(3):
        for d in doses:
            d.fill_(-2.0)
        sched.run(graph=True)
        t.cuda.synchronize()
        for d, r in zip(doses, refs):
            assert t.equal(d, r), rep
        # eager work of the program between replays
        engine.photon_dose(cases[0][0], cases[0][1], np.zeros(3), cases[0][2], 100.0, 10.0, 0.0, 0.0, pm,
                           square_field_mlc(pm, 5.0, 5.0))
        t.cuda.synchronize()
    sched.close()


def test_scheduler_segments_share_ray_trace(photon_model):
    """Control points of one case with the same isocentre, gantry and couch
    (different MLC segments, collimator angles and MU) reuse the first one's
    ray trace inside the batch; the case dose still equals the job-ordered sum
    of independent single-control-point doses bit for bit, eager and as a
    replayed graph, including after a control point with another gantry."""
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import square_field_mlc
    from dosecuda_amd.scheduler import Scheduler
    pm = photon_model
    rng = np.random.default_rng(23)
    cases = _photon_cases(t, rng, n_cases=3)
    sched = Scheduler(0, streams=2)
    doses, refs = [], []
    for rho, s, iso, _ in cases:
        d = t.zeros_like(rho)
        cid = sched.add_case(rho, s, np.zeros(3), d)
        ref = t.zeros_like(rho)
        for ga, segs in ((30.0, 3), (200.0, 2), (30.0, 2)):
            for k in range(segs):
                mlc = square_field_mlc(pm, float(rng.uniform(2, 8)), float(rng.uniform(2, 8)))
                ca, mu = float(rng.uniform(0, 90)), float(rng.uniform(20, 120))
                sched.add_photon(cid, iso, mu, ga, ca, 0.0, pm, mlc)
                ref += engine.photon_dose(rho, s, np.zeros(3), iso, mu, ga, ca, 0.0, pm, mlc)
        doses.append(d)
        refs.append(ref)
    print(f"[case] scheduler shared ray trace, {len(cases)} cases x 7 control points", flush=True)
    sched.run(graph=False)
    t.cuda.synchronize()
    for d, r in zip(doses, refs):
        assert float(r.max()) > 0.0
        assert t.equal(d, r)
    for rep in range(2):
        for d in doses:
            d.fill_(-1.0)
        sched.run(graph=True)
        t.cuda.synchronize()
        for d, r in zip(doses, refs):
            assert t.equal(d, r), rep
    sched.close()


def test_scheduler_proton_wet_dose_and_influence(proton_model):
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.scheduler import Scheduler
    m = proton_model
    rng = np.random.default_rng(5)
    shape, s = (30, 40, 36), 2.5
    rlsp = t.ones(shape, device="cuda")
    rlsp[:, 15:20, :] = 1.6
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    ids = np.sort(rng.choice(np.arange(20, 40), 3, replace=False))
    fields = []
    for g in (0.0, 90.0):
        xy = rng.uniform(-15.0, 15.0, (60, 2))
        spots = np.column_stack([xy, np.ones(60), np.repeat(ids, 20)]).astype(np.float32)
        fields.append((g, spots))
    print("[case] scheduler proton", flush=True)
    sched = Scheduler(0, streams=2)
    dose = t.full_like(rlsp, -1.0)
    cid = sched.add_case(rlsp, s, np.zeros(3), dose)
    ref = t.zeros_like(rlsp)
    wets, plans, refs_csr, outs_csr = [], [], [], []
    for g, spots in fields:
        w = t.full_like(rlsp, -1.0)
        plan = engine.ProtonPlan(0, shape, s, np.zeros(3), iso, g, 0.0, m, spots)
        sched.add_proton_wet(cid, iso, g, 0.0, m, w)
        sched.add_proton_dose(cid, plan, w)
        wr = engine.proton_wet(rlsp, s, np.zeros(3), iso, g, 0.0, m.VSADX, m.VSADY)
        ref += plan.dose(wr)
        wets.append((w, wr))
        plans.append(plan)
    # influence matrices of the same fields, in a case of their own
    cinf = sched.add_case(rlsp, s, np.zeros(3), None)
    for g, spots in fields:
        wr = engine.proton_wet(rlsp, s, np.zeros(3), iso, g, 0.0, m.VSADX, m.VSADY)
        f = engine.InfluenceMatrix(0, shape, s, np.zeros(3), iso, g, 0.0, m, spots)
        nnz = f.count(wr)
        col0 = t.empty(nnz, dtype=t.int32, device="cuda")
        val0 = t.empty(nnz, dtype=t.float32, device="cuda")
        f.fill(wr, col0, val0)
        col = t.full((nnz,), -1, dtype=t.int32, device="cuda")
        val = t.full((nnz,), -1.0, dtype=t.float32, device="cuda")
        sched.add_influence_fill(cinf, f, wr, col, val)
        refs_csr.append((col0, val0, f, wr))
        outs_csr.append((col, val))
    for graph in (False, True, True):
        dose.fill_(-3.0)
        for (col, val) in outs_csr:
            col.fill_(-1)
            val.fill_(-1.0)
        sched.run(graph=graph)
        t.cuda.synchronize()
        for w, wr in wets:
            assert t.equal(w, wr)
        assert t.equal(dose, ref)
        for (col, val), (col0, val0, _, _) in zip(outs_csr, refs_csr):
            assert t.equal(col, col0) and t.equal(val, val0)
    sched.close()


def test_scheduler_surrogate_photon(photon_model):
    t = _torch()
    from dosecuda_amd import _lib, engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    from dosecuda_amd.scheduler import Scheduler
    pm = photon_model
    net = S.SurrogateNet().cuda()
    net.out_scale.fill_(1.5)
    native = S.NativeNet(net)
    rng = np.random.default_rng(3)
    cases = _photon_cases(t, rng, n_cases=3)
    print("[case] scheduler surrogate", flush=True)
    sched = Scheduler(0, streams=2)
    doses, refs = [], []
    for rho, s, iso, beams in cases:
        d = t.empty_like(rho)
        cid = sched.add_case(rho, s, np.zeros(3), d)
        ref = t.zeros_like(rho)
        for ga, fs in beams:
            mlc = square_field_mlc(pm, fs, fs)
            sched.add_photon(cid, iso, 100.0, ga, 0.0, 0.0, pm, mlc, surrogate=native, stride=4, tol=0.02, dilate=1)
            one = t.empty_like(rho)
            keep: list = []
            m = engine.photon_args(pm, keep)
            rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
            _lib.check(_lib.load().dh_photon_dose_surrogate_dev(
                0, native.handle, _lib.grid(rho.shape, s, np.zeros(3)), _lib.photon_cp(iso, 100.0, ga, 0.0, 0.0),
                m, _lib.ptr(rows), rows.shape[0], _lib.ptr(rho), _lib.ptr(one), None, None, 4, 0.02, 1, S._st()))
            ref += one
        doses.append(d)
        refs.append(ref)
    for graph in (False, True):
        sched.run(graph=graph)
        t.cuda.synchronize()
        for d, r in zip(doses, refs):
            assert t.equal(d, r)
    sched.close()


def test_scheduler_surrogate_groups(photon_model):
    """Eleven checked-surrogate control points of one case (a group of 8 and
    one of 3, runs of equal gantry angles sharing the depth, scaled doses)
    equal the scaled sum of single-point surrogate doses bit for bit, eager
    and as a graph."""
    t = _torch()
    from dosecuda_amd import _lib, engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    from dosecuda_amd.scheduler import Scheduler
    pm = photon_model
    net = S.SurrogateNet().cuda()
    net.out_scale.fill_(1.5)
    native = S.NativeNet(net)
    rng = np.random.default_rng(7)
    rho, s, iso, _ = _photon_cases(t, rng, n_cases=1)[0]
    gantries = [0.0, 0.0, 0.0, 40.0, 40.0, 90.0, 90.0, 90.0, 200.0, 200.0, 300.0]
    print(f"[case] scheduler surrogate groups of {len(gantries)} control points", flush=True)
    sched = Scheduler(0, streams=1)
    d = t.empty_like(rho)
    cid = sched.add_case(rho, s, np.zeros(3), d)
    ref = t.zeros_like(rho)
    for q, ga in enumerate(gantries):
        fs = 4.0 + (q % 4)
        scale = float(np.float32(0.5 + 0.25 * q))
        mlc = square_field_mlc(pm, fs, fs + 1.0)
        sched.add_photon(cid, iso, 100.0, ga, 0.0, 0.0, pm, mlc, surrogate=native, stride=4, tol=0.02, dilate=1,
                         scale=scale)
        one = t.empty_like(rho)
        keep: list = []
        m = engine.photon_args(pm, keep)
        rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
        _lib.check(_lib.load().dh_photon_dose_surrogate_dev(
            0, native.handle, _lib.grid(rho.shape, s, np.zeros(3)), _lib.photon_cp(iso, 100.0, ga, 0.0, 0.0),
            m, _lib.ptr(rows), rows.shape[0], _lib.ptr(rho), _lib.ptr(one), None, None, 4, 0.02, 1, S._st()))
        ref = ref + one * scale
    for graph in (False, True, True):
        sched.run(graph=graph)
        t.cuda.synchronize()
        assert t.equal(d, ref)
    sched.close()
