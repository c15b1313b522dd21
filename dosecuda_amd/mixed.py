"""The mixed batched workload on one GPU: IMPT cases as influence matrices
(WET per field, counted once at set-up (timed: count_s), filled in the batch) and IMRT cases as the
checked surrogate CCCS per MLC segment (or the plain CCCS), all through one
scheduler whose batch is replayed as a single HIP graph.

Error accounting per case: an IMRT case's surrogate plan dose against the
same plan through the plain CCCS (max |diff| / Dmax and global 2%/2 mm
gamma); an IMPT case's influence matrix applied to spot weights against
the dense pencil-beam dose with those weights (max |diff| / Dmax, bounded by
the 1e-4-of-row-peak truncation).  The plain paths themselves are checked
against the FP64 oracle by the parity tests.

No reference counterpart.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import engine
from .beam_models import density_from_hu, rlsp_from_hu
from .phantoms import MixedCase, mixed_iso, voxelize_mixed
from .scheduler import Scheduler


def _material(hu: torch.Tensor, conv) -> torch.Tensor:
    """HU -> RLSP / density through the table conversion; the phantoms hold
    a few distinct HU values, converted once each."""
    vals = torch.unique(hu)
    out = torch.empty_like(hu)
    conv_vals = conv(vals.cpu().numpy().astype(np.float32))
    for v, c in zip(vals.tolist(), np.asarray(conv_vals, np.float32).tolist()):
        out[hu == v] = c
    return out


class MixedWorkload:
    def __init__(self, cases: list, pmodel, xmodel, surrogate=None, streams=4, stride=3, tol=0.01, dilate=1,
                 log=None):
        self.cases, self.pmodel, self.xmodel = cases, pmodel, xmodel
        self.surrogate = surrogate
        self.sched = Scheduler(0, streams=streams)
        self.items = []
        self.nnz = 0
        self.n_spots = 0
        self.n_cps = 0
        self.voxel_cps = 0
        self.count_s = 0.0                 # wall time of the IMPT fields' peak/count passes
        t0 = time.time()
        for ci, c in enumerate(cases):
            hu = voxelize_mixed(c, "cuda")
            iso = mixed_iso(c)
            org = np.zeros(3, np.float32)
            if c.kind == "impt":
                rlsp = _material(hu, rlsp_from_hu)
                cid = self.sched.add_case(rlsp, c.spacing, org, None)
                fl = []
                for g, spots in c.fields:
                    wet = engine.proton_wet(rlsp, c.spacing, org, iso, g, 0.0, pmodel.VSADX, pmodel.VSADY)
                    f = engine.InfluenceMatrix(0, c.shape, c.spacing, org, iso, g, 0.0, pmodel, spots)
                    torch.cuda.synchronize()
                    tc = time.perf_counter()
                    nnz = f.count(wet)           # bounds, peak and count passes (host-synchronised)
                    self.count_s += time.perf_counter() - tc
                    col = torch.empty(max(nnz, 1), dtype=torch.int32, device="cuda")
                    val = torch.empty(max(nnz, 1), dtype=torch.float32, device="cuda")
                    self.sched.add_proton_wet(cid, iso, g, 0.0, pmodel, wet)
                    self.sched.add_influence_fill(cid, f, wet, col, val)
                    fl.append(dict(gantry=g, spots=spots, wet=wet, f=f, col=col, val=val, nnz=nnz))
                    self.nnz += nnz
                    self.n_spots += len(spots)
                self.items.append(dict(kind="impt", case=c, rlsp=rlsp, iso=iso, fields=fl))
            else:
                rho = _material(hu, density_from_hu)
                dose = torch.empty_like(rho)
                cid = self.sched.add_case(rho, c.spacing, org, dose)
                for g, mu, mlc in c.cps:
                    self.sched.add_photon(cid, iso, mu, g, 0.0, 0.0, xmodel, mlc, surrogate=surrogate,
                                          stride=stride, tol=tol, dilate=dilate)
                    self.n_cps += 1
                    self.voxel_cps += int(np.prod(c.shape))
                self.items.append(dict(kind="imrt", case=c, rho=rho, iso=iso, dose=dose))
            if log and (ci + 1) % 8 == 0:
                log(f"mixed: set up {ci + 1}/{len(cases)} cases ({time.time() - t0:.1f} s)")
        torch.cuda.synchronize()
        self.setup_s = time.time() - t0

    def run(self, graph=True):
        self.sched.run(graph=graph)

    def close(self):
        self.sched.close()
        for it in self.items:
            for f in it.get("fields", []):
                f["f"].close()

    def counts(self):
        return dict(cases=len(self.cases), impt_cases=sum(1 for c in self.cases if c.kind == "impt"),
                    imrt_cases=sum(1 for c in self.cases if c.kind == "imrt"), spots=self.n_spots, nnz=self.nnz,
                    control_points=self.n_cps, voxel_cps=self.voxel_cps)

    def errors(self, log=None, gamma_cases=None):
        """Per-case errors (see module doc).  Runs the plain photon plans in
        a scheduler of their own and the dense proton doses."""
        from .surrogate import gamma
        out = []
        plain = Scheduler(0, streams=self.sched.n_streams)
        refs = []
        for it in self.items:
            if it["kind"] == "imrt":
                c = it["case"]
                ref = torch.empty_like(it["rho"])
                cid = plain.add_case(it["rho"], c.spacing, np.zeros(3), ref)
                for g, mu, mlc in c.cps:
                    plain.add_photon(cid, it["iso"], mu, g, 0.0, 0.0, self.xmodel, mlc)
                refs.append(ref)
        plain.run(graph=False)
        torch.cuda.synchronize()
        plain.close()
        k = 0
        rng = np.random.default_rng(7)
        for i, it in enumerate(self.items):
            c = it["case"]
            if it["kind"] == "imrt":
                ref = refs[k]
                k += 1
                d = it["dose"]
                dm = float(ref.max())
                e = dict(case=i, kind="imrt", shape=c.shape, spacing=c.spacing, cps=len(c.cps),
                         max_err=float((d - ref).abs().max()) / dm,
                         rms_err=float((d - ref).pow(2).mean().sqrt()) / dm)
                if gamma_cases is None or i in gamma_cases:
                    e["gamma"] = gamma(d, ref, c.spacing)[0]
            else:
                worst = 0.0
                for fl in it["fields"]:
                    spots = fl["spots"].copy()
                    w = rng.uniform(0.5, 2.0, len(spots)).astype(np.float32)
                    spots[:, 2] = w
                    plan = engine.ProtonPlan(0, c.shape, c.spacing, np.zeros(3), it["iso"], fl["gantry"], 0.0,
                                             self.pmodel, spots)
                    dense = plan.dose(fl["wet"])
                    row_spot = fl["f"].info()[0]
                    rp = torch.from_numpy(np.asarray(fl["f"].row_ptr, np.int64)).cuda()
                    counts = rp[1:] - rp[:-1]
                    wt = torch.from_numpy(w[np.asarray(row_spot)]).cuda()
                    rows = torch.repeat_interleave(torch.arange(len(counts), device="cuda"), counts)
                    acc = torch.zeros(int(np.prod(c.shape)), dtype=torch.float64, device="cuda")
                    nz = int(rp[-1] - rp[0])
                    acc.index_add_(0, fl["col"][:nz].long(), (fl["val"][:nz].double() * wt[rows].double()))
                    dm = float(dense.max())
                    worst = max(worst, float((acc.view(c.shape).float() - dense).abs().max()) / dm)
                e = dict(case=i, kind="impt", shape=c.shape, spacing=c.spacing, fields=len(it["fields"]),
                         nnz=sum(f["nnz"] for f in it["fields"]), max_err=worst)
            out.append(e)
            if log:
                log(str(e))
        return out
