"""GPU: the mixed workload at small sizes.  The batch replayed as a HIP
graph equals the eager batch; influence matrices applied to spot weights
match the dense pencil-beam dose within the truncation; one IMPT field and
one IMRT plan (plain CCCS) match the FP64 oracle within 1e-4 of Dmax; with
the trained surrogate, every IMRT case passes 2%/2 mm gamma >= 99% against
the plain plan."""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _small_cases(pm, xm, seed=21, n=4):
    from dosecuda_amd import phantoms
    return phantoms.mixed_cases(pm, xm, seed=seed, n_cases=n, sides=(28, 44), spacing=(2.0, 3.0),
                                spots=(60, 120), fields=(1, 2), beams=(2, 3), segments=(1, 3))


def test_mixed_graph_equals_eager_and_oracle(proton_model, photon_model):
    t = _torch()
    from dosecuda_amd import engine, mixed
    cases = _small_cases(proton_model, photon_model)
    print(f"[case] mixed small: {[(c.kind, c.shape) for c in cases]}", flush=True)
    W = mixed.MixedWorkload(cases, proton_model, photon_model, surrogate=None, streams=3)
    W.run(graph=False)
    t.cuda.synchronize()
    eager = [(it["dose"].clone() if it["kind"] == "imrt" else [f["val"].clone() for f in it["fields"]])
             for it in W.items]
    for _ in range(2):
        W.run(graph=True)
        t.cuda.synchronize()
        for e, it in zip(eager, W.items):
            if it["kind"] == "imrt":
                assert t.equal(e, it["dose"])
            else:
                assert all(t.equal(a, f["val"]) for a, f in zip(e, it["fields"]))
    errs = W.errors()
    for e in errs:
        print(f"[parity] mixed {e}", flush=True)
        if e["kind"] == "impt":
            assert e["max_err"] < 2e-3           # truncation at 1e-4 of each row peak
        else:
            assert e["max_err"] == 0.0           # plain scheduler vs plain scheduler
    # oracle: first IMRT case (plain CCCS plan) and first IMPT field
    im = next(it for it in W.items if it["kind"] == "imrt")
    c = im["case"]
    rho = im["rho"].cpu().numpy()
    ref = np.zeros(c.shape)
    for g, mu, mlc in c.cps:
        ref += O.photon_dose(rho, c.spacing, np.zeros(3), im["iso"], mu, g, 0.0, 0.0, photon_model, mlc)
    d = im["dose"].cpu().numpy()
    err = np.abs(d - ref).max() / ref.max()
    print(f"[parity] mixed imrt plan vs oracle max {err:.2e}", flush=True)
    assert err <= 1e-4
    ip = next(it for it in W.items if it["kind"] == "impt")
    c, fl = ip["case"], ip["fields"][0]
    wet = fl["wet"].cpu().numpy()
    m = proton_model
    oref, _ = O.proton_dose(wet, c.spacing, np.zeros(3), ip["iso"], fl["gantry"], 0.0, m.VSADX, m.VSADY, fl["spots"],
                            m.divergence_params, m.lut_depths, m.lut_sigmas, m.lut_idds)
    plan = engine.ProtonPlan(0, c.shape, c.spacing, np.zeros(3), ip["iso"], fl["gantry"], 0.0, m, fl["spots"])
    got = plan.dose(fl["wet"]).cpu().numpy()
    err = np.abs(got - oref).max() / oref.max()
    print(f"[parity] mixed impt field vs oracle max {err:.2e}", flush=True)
    assert err <= 1e-4
    W.close()


def test_mixed_with_trained_surrogate(proton_model, photon_model):
    t = _torch()
    from dosecuda_amd import mixed, surrogate as S
    if not os.path.exists(S.WEIGHTS):
        pytest.skip("no trained surrogate weights")
    net = S.load(S.WEIGHTS)
    cases = [c for c in _small_cases(proton_model, photon_model, seed=33, n=6) if c.kind == "imrt"]
    print(f"[case] mixed surrogate: {[c.shape for c in cases]}", flush=True)
    # the strict check setting (26-neighbour dilation); the bench reports the default one
    W = mixed.MixedWorkload(cases, proton_model, photon_model, surrogate=S.NativeNet(net), streams=2, tol=0.01,
                            dilate=2)
    W.run(graph=True)
    t.cuda.synchronize()
    for e in W.errors():
        print(f"[parity] mixed surrogate {e}", flush=True)
        assert e["gamma"] >= 0.99
    W.close()
