"""GPU: the batched CSR dose-influence matrix against the FP64 oracle.

Each row must equal the oracle's single-spot (MU 1) dose of that spot above
the row's truncation threshold (1e-4 of the row peak): same support (up to
entries within 1e-4 relative of the threshold), values within 1e-5 of the
row peak, ascending columns.  At full size: every slot written, sorted rows,
and sum_j MU_j row_j equal to the dense pencil-beam dose within the summed
truncation thresholds.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
F = np.float32
REL = 1e-4


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _case(shape=(40, 48, 44), s=2.5, seed=41):
    rng = np.random.default_rng(seed)
    rho = np.full(shape, 1.00653, F)
    k, j, i = np.ogrid[:shape[0], :shape[1], :shape[2]]
    for _ in range(3):
        c = rng.uniform(0.25, 0.75, 3) * np.array(shape)
        r = rng.uniform(3, 7)
        m = (k - c[0]) ** 2 + (j - c[1]) ** 2 + (i - c[2]) ** 2 <= r * r
        rho[np.broadcast_to(m, shape)] = rng.choice([0.26, 1.56, 1.75])
    iso = np.array([(shape[2] - 1) * s / 2, (shape[1] - 1) * s / 2, (shape[0] - 1) * s / 2], F)
    return rho, iso, s


@pytest.mark.parametrize("gantry, S, C", [(0.0, 8, 4), (120.0, 16, 3), (250.0, 8, 1), (60.0, 4, 2)])
def test_influence_rows_match_oracle(proton_model, gantry, S, C):
    t = _torch()
    from dosecuda_amd import engine
    m = proton_model
    rho, iso, s = _case()
    rng = np.random.default_rng(int(gantry) + 5)
    n = 60
    ids = rng.choice([8, 14, 22], n)
    spots = np.column_stack([rng.uniform(-25, 25, n), rng.uniform(-20, 20, n), np.ones(n), ids]).astype(F)
    print(f"[case] influence gantry={gantry} S={S} C={C} spots={n}", flush=True)
    d = t.from_numpy(rho).cuda()
    wet = engine.proton_wet(d, s, np.zeros(3), iso, gantry, 0.0, m.VSADX, m.VSADY)
    wet32 = wet.cpu().numpy()
    infl = engine.InfluenceMatrix(0, rho.shape, s, np.zeros(3), iso, gantry, 0.0, m, spots,
                                  spots_per_block=S, row_chunks=C, rel=REL)
    row_ptr, col, val = infl.build(wet)
    row_ptr, col, val = row_ptr.cpu().numpy(), col.cpu().numpy().view(np.uint32), val.cpu().numpy()
    row_spot, row_peak = infl.info()
    assert row_ptr[0] == 0 and row_ptr[-1] == col.size and np.all(np.diff(row_ptr) >= 0)
    g = O.geom(rho.shape, s, np.zeros(3), iso, gantry, 0.0, 0.0, (F(m.VSADX) + F(m.VSADY)) / F(2.0))
    worst_v, worst_p, n_flip = 0.0, 0.0, 0
    for q in range(n):
        sp = spots[row_spot[q]:row_spot[q] + 1].copy()
        sp[0, 2] = 1.0
        ref, _ = O.pencil(g, m.VSADX, m.VSADY, wet32, sp, m.divergence_params, m.lut_depths,
                          m.lut_sigmas, m.lut_idds)
        ref = ref.ravel()
        pk = ref.max()
        worst_p = max(worst_p, abs(row_peak[q] - pk) / pk)
        c = col[row_ptr[q]:row_ptr[q + 1]].astype(np.int64)
        v = val[row_ptr[q]:row_ptr[q + 1]]
        assert np.all(np.diff(c) > 0), "columns must ascend"
        thr = REL * pk
        want = np.nonzero(ref >= thr * (1 + 1e-4))[0]
        assert np.all(np.isin(want, c)), f"row {q}: missing entries"
        assert np.all(ref[c] >= thr * (1 - 1e-4)), f"row {q}: entries below threshold"
        n_flip += int(np.count_nonzero(~np.isin(c, want)))
        worst_v = max(worst_v, float(np.abs(v - ref[c]).max() / pk) if c.size else 0.0)
    print(f"[parity] influence rows: max |val - oracle| / peak {worst_v:.3e}, peak rel err "
          f"{worst_p:.3e}, nnz {col.size}, entries at the threshold margin {n_flip}", flush=True)
    assert worst_v <= 1e-5 and worst_p <= 1e-5
    infl.close()


def test_influence_shared_csr_equals_per_field(proton_model):
    """Two fields counted into one CSR (the second starting at the first's
    entry count, `base`) and filled: every slot written, and each field's
    rows equal, bit for bit, the CSR that field builds alone."""
    t = _torch()
    from dosecuda_amd import engine
    m = proton_model
    rho, iso, s = _case(seed=43)
    d = t.from_numpy(rho).cuda()
    rng = np.random.default_rng(9)
    fields = []
    for gantry in (30.0, 200.0):
        n = 50
        spots = np.column_stack([rng.uniform(-25, 25, n), rng.uniform(-20, 20, n), np.ones(n),
                                 rng.choice([10, 18], n)]).astype(F)
        wet = engine.proton_wet(d, s, np.zeros(3), iso, gantry, 0.0, m.VSADX, m.VSADY)
        fields.append((gantry, spots, wet))
    print("[case] influence shared CSR of 2 fields", flush=True)
    alone = []
    for gantry, spots, wet in fields:
        f = engine.InfluenceMatrix(0, rho.shape, s, np.zeros(3), iso, gantry, 0.0, m, spots)
        rp, col, val = f.build(wet)
        alone.append((rp.cpu().numpy(), col.clone(), val.clone()))
        f.close()
    hs = [engine.InfluenceMatrix(0, rho.shape, s, np.zeros(3), iso, g, 0.0, m, sp) for g, sp, _ in fields]
    base = 0
    for f, (_, _, wet) in zip(hs, fields):
        base += f.count(wet, base)
    col = t.full((base,), -1, dtype=t.int32, device="cuda")
    val = t.zeros((base,), dtype=t.float32, device="cuda")
    for f, (_, _, wet) in zip(hs, fields):
        f.fill(wet, col, val)
    t.cuda.synchronize()
    assert int((col == -1).sum().item()) == 0, "unwritten slots"
    for f, (rp1, c1, v1) in zip(hs, alone):
        rp = np.asarray(f.row_ptr)
        assert np.array_equal(rp - rp[0], rp1)
        a, b = int(rp[0]), int(rp[-1])
        assert t.equal(col[a:b], c1) and t.equal(val[a:b], v1)
        f.close()
    print(f"[parity] shared CSR: {base} entries, both fields equal their own CSR", flush=True)


def test_influence_count_many_equals_sequential(proton_model):
    """Five fields' footprint WET and counts overlapped on the library's four
    streams (two fields share one) and their fills overlapped on side streams
    give the WET, entry counts, row offsets and filled CSR of the single-field
    calls in order, bit for bit; a second call
    on the same handles (their footprints now from the first call's
    thresholds, so the WET outside them differs) gives the same CSR again."""
    t = _torch()
    from dosecuda_amd import engine
    m = proton_model
    rho, iso, s = _case(seed=44)
    d = t.from_numpy(rho).cuda()
    rng = np.random.default_rng(11)
    specs = []
    for gantry in (0.0, 75.0, 150.0, 225.0, 300.0):
        n = 40
        spots = np.column_stack([rng.uniform(-25, 25, n), rng.uniform(-20, 20, n), np.ones(n),
                                 rng.choice([8, 14, 20], n)]).astype(F)
        specs.append((gantry, spots))
    print("[case] influence count_many of 5 fields on 4 streams", flush=True)
    mk = lambda: [engine.InfluenceMatrix(0, rho.shape, s, np.zeros(3), iso, g, 0.0, m, sp) for g, sp in specs]

    def fill_all(hs, wets, total, overlapped=False):
        col = t.full((total,), -1, dtype=t.int32, device="cuda")
        val = t.zeros((total,), dtype=t.float32, device="cuda")
        if overlapped:
            engine.InfluenceMatrix.fill_many(hs, wets, col, val)
        else:
            for f, w in zip(hs, wets):
                f.fill(w, col, val)
        t.cuda.synchronize()
        return col, val

    seq = mk()
    wets_seq = [f.wet(d) for f in seq]
    base, counts = 7, []
    for f, w in zip(seq, wets_seq):
        counts.append(f.count(w, base + sum(counts)))
    col_s, val_s = fill_all(seq, wets_seq, base + sum(counts))
    many = mk()
    for rep in range(2):
        wets_m = [t.empty_like(d) for _ in many]
        got = engine.InfluenceMatrix.count_many(many, d, wets_m, base=base)
        t.cuda.synchronize()
        assert got == counts, (got, counts)
        for f0, f1, w0, w1 in zip(seq, many, wets_seq, wets_m):
            assert rep == 1 or t.equal(w0, w1)
            assert np.array_equal(np.asarray(f0.row_ptr), np.asarray(f1.row_ptr))
        col_m, val_m = fill_all(many, wets_m, base + sum(got), overlapped=True)
        assert int((col_m[base:] == -1).sum().item()) == 0, "unwritten slots"
        assert t.equal(col_m[base:], col_s[base:]) and t.equal(val_m[base:], val_s[base:])
    for f in seq + many:
        f.close()
    print(f"[parity] count_many: {sum(counts)} entries in 5 fields equal the sequential build", flush=True)


def _csr_equal(t, a, b):
    return np.array_equal(a[0], b[0]) and t.equal(a[1], b[1]) and t.equal(a[2], b[2])


@pytest.mark.parametrize("full_size", [False, True])
def test_influence_footprint_wet_gives_the_same_matrix(proton_model, full_size):
    """The WET computed only on the matrix's footprint (dh_influence_wet)
    equals the full-grid WET bit for bit inside the footprint, and the CSR
    built from it equals the CSR of the full-grid WET bit for bit.  With the
    margin forced to zero some spot reaches past the footprint: count()
    falls back to the full grid and the CSR is still the same."""
    t = _torch()
    from dosecuda_amd import engine, phantoms
    m = proton_model
    if full_size:
        case = phantoms.spheres_case(seed=1003)
        rho, s, iso, org = phantoms.to_rlsp(case), case.spacing, case.iso, case.origin
        gantry, spots = phantoms.influence_fields(m, seed=1003)[1]
    else:
        rho, iso, s = _case(seed=45)
        org = np.zeros(3)
        rng = np.random.default_rng(11)
        n = 60
        spots = np.column_stack([rng.uniform(-15, 15, n), rng.uniform(-12, 12, n), np.ones(n),
                                 rng.choice([8, 14, 22], n)]).astype(F)
        gantry = 75.0
    print(f"[case] influence footprint full_size={full_size} gantry={gantry} spots={len(spots)}", flush=True)
    d = t.from_numpy(rho).cuda()
    full = engine.proton_wet(d, s, org, iso, gantry, 0.0, m.VSADX, m.VSADY)
    f_ref = engine.InfluenceMatrix(0, rho.shape, s, org, iso, gantry, 0.0, m, spots)
    rp, col, val = f_ref.build(full)
    ref = (rp.cpu().numpy(), col.clone(), val.clone())
    f_ref.close()
    for scale, expect_fallback in ((1.0, False), (0.0, True)):
        f = engine.InfluenceMatrix(0, rho.shape, s, org, iso, gantry, 0.0, m, spots)
        w = f.wet(d, margin_scale=scale)
        inside = w < 1e29
        frac = float(inside.float().mean())
        assert t.equal(w[inside], full[inside])
        got = f.build(w)
        box, active, fb = f.footprint()
        print(f"[parity] footprint scale {scale}: box {np.round(box, 1).tolist()}, {frac:.3f} of voxels, "
              f"fallbacks {fb}", flush=True)
        assert (fb > 0) == expect_fallback and active == (not expect_fallback)
        assert _csr_equal(t, (got[0].cpu().numpy(), got[1], got[2]), ref)
        if expect_fallback:
            assert t.equal(w, full)          # the fallback rewrote the whole WET
        f.close()


def test_influence_full_size_properties(proton_model):
    """Batched IMPT influence matrix at full size: (160, 256, 256) @ 2 mm with
    seeded spheres, 3 fields at gantry 0/120/240, ~15k spots over 25 layers."""
    t = _torch()
    from dosecuda_amd import engine, phantoms
    m = proton_model
    case = phantoms.spheres_case(seed=1003)
    rlsp = phantoms.to_rlsp(case)
    d = t.from_numpy(rlsp).cuda()
    fields = phantoms.influence_fields(m, seed=1003)
    for gantry, spots in fields:
        print(f"[case] influence full gantry={gantry} spots={len(spots)}", flush=True)
        wet = engine.proton_wet(d, case.spacing, case.origin, case.iso, gantry, 0.0, m.VSADX, m.VSADY)
        infl = engine.InfluenceMatrix(0, rlsp.shape, case.spacing, case.origin, case.iso, gantry,
                                      0.0, m, spots)
        nnz = infl.count(wet)
        col = t.full((nnz,), -1, dtype=t.int32, device="cuda")
        val = t.zeros((nnz,), dtype=t.float32, device="cuda")
        infl.fill(wet, col, val)
        rp = t.from_numpy(infl.row_ptr).cuda()
        assert int((col == -1).sum().item()) == 0, "unwritten slots"
        # rows ascend: a column may only decrease where a new row starts
        dec = (col[1:].long() - col[:-1].long()) <= 0
        starts = t.zeros(nnz, dtype=t.bool, device="cuda")
        inner = rp[1:-1]
        inner = inner[(inner > 0) & (inner < nnz)]
        starts[inner] = True
        assert int((dec & ~starts[1:]).sum().item()) == 0, "unsorted row"
        # sum_j mu_j row_j vs the dense pencil sum (mu = 1)
        spot_of_row, peak = infl.info()
        rows = t.repeat_interleave(t.arange(len(spots), device="cuda"), rp[1:] - rp[:-1])
        recon = t.zeros(rlsp.size, dtype=t.float64, device="cuda")
        recon.index_add_(0, col.long(), val.double())
        plan = engine.ProtonPlan(0, rlsp.shape, case.spacing, case.origin, case.iso, gantry, 0.0, m,
                                 spots)
        dense = plan.dose(wet).double().view(-1)
        bound = float(REL * peak.astype(np.float64).sum())
        err = float((dense - recon).abs().max().item())
        dmax = float(dense.max().item())
        print(f"[parity] influence full gantry={gantry}: nnz {nnz}, max |dense - sum rows| {err:.4g} "
              f"(truncation bound {bound:.4g}, Dmax {dmax:.4g})", flush=True)
        assert err <= bound + 1e-5 * dmax
        del rows
        plan.close()
        infl.close()
