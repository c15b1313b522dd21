"""GPU: the surrogate's BF16 MFMA convolution kernels against FP32 torch
convolutions on the CPU (same bf16-rounded operands), the gamma index
against a brute-force NumPy gamma and its known answers, the voxel-list
CCCS, and a short end-to-end training + fallback run."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _ref_conv(x_bf, w32, b, ksize, act, slope):
    """fp32 CPU reference on channels-last tensors with bf16-rounded inputs."""
    import torch
    import torch.nn.functional as F
    x = x_bf.float().cpu().permute(0, 4, 1, 2, 3)
    w = w32.to(torch.bfloat16).float().cpu()                # [taps, cin, cout]
    taps, cin, cout = w.shape
    wk = w.view(ksize, ksize, ksize, cin, cout).permute(4, 3, 0, 1, 2).contiguous()
    y = F.conv3d(x, wk, b.cpu() if b is not None else None, padding=ksize // 2)
    if act:
        y = torch.where(y > 0, y, y * slope)
    return y.permute(0, 2, 3, 4, 1)


@pytest.mark.parametrize("shape, cin, cout, ksize, act", [
    ((2, 6, 7, 9), 32, 48, 3, True),
    ((1, 8, 8, 8), 16, 32, 3, False),
    ((1, 5, 6, 11), 48, 16, 1, False),
    ((1, 20, 30, 60), 32, 16, 3, True),      # large: LDS brick kernel, partial x and y tiles
    ((2, 6, 12, 64), 48, 48, 3, False),      # large, batch of 2, 64 padded output channels
])
def test_conv3d_forward(shape, cin, cout, ksize, act):
    t = _torch()
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(7)
    x = t.randn(shape + (cin,), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((ksize ** 3, cin, cout), generator=g) * 0.1).cuda()
    b = (t.randn(cout, generator=g) * 0.1).cuda()
    print(f"[case] conv3d fwd {shape} {cin}->{cout} k{ksize}", flush=True)
    y, y32 = S.conv3d(x, S.prep_weights(w, 0), b, cout, ksize, act, want_bf16=True, want_f32=True)
    ref = _ref_conv(x, w, b, ksize, act, S.SLOPE)
    err = float((y32.cpu() - ref).abs().max() / ref.abs().max())
    print(f"[parity] conv3d fwd rel err {err:.2e}", flush=True)
    assert err < 1e-5
    assert float((y.float().cpu() - ref).abs().max() / ref.abs().max()) < 1e-2   # bf16 output


def test_conv3d_backward():
    t = _torch()
    from dosecuda_amd import surrogate as S
    import torch.nn.functional as F
    g = t.Generator().manual_seed(11)
    shape, cin, cout = (1, 6, 5, 7), 32, 32
    x = t.randn(shape + (cin,), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((27, cin, cout), generator=g) * 0.1).cuda().requires_grad_(True)
    b = t.zeros(cout).cuda().requires_grad_(True)
    xg = x.clone().requires_grad_(True)
    print("[case] conv3d backward", flush=True)
    y = S._Conv.apply(xg, w, b, 3, True, False)
    gy = t.randn(y.shape, generator=g).to(t.bfloat16).cuda()
    y.backward(gy)
    # fp32 CPU reference with the same bf16-rounded operands
    xr = x.float().cpu().requires_grad_(True)
    wr = w.detach().to(t.bfloat16).float().cpu().requires_grad_(True)
    wk = wr.view(3, 3, 3, cin, cout).permute(4, 3, 0, 1, 2)
    yr = F.conv3d(xr.permute(0, 4, 1, 2, 3), wk, None, padding=1)
    yr = t.where(yr > 0, yr, yr * S.SLOPE).permute(0, 2, 3, 4, 1)
    yr.backward(gy.float().cpu())
    ex = float((xg.grad.float().cpu() - xr.grad).abs().max() / xr.grad.abs().max())
    ew = float((w.grad.cpu() - wr.grad).abs().max() / wr.grad.abs().max())
    print(f"[parity] conv3d dgrad rel err {ex:.2e}, wgrad rel err {ew:.2e}", flush=True)
    assert ex < 1e-2 and ew < 1e-2


@pytest.mark.parametrize("nco, ks", [(1, 1), (1, 2), (1, 4), (1, 8), (1, 16), (2, 1), (2, 2), (2, 4),
                                     (2, 8), (4, 1), (4, 2), (4, 4), (1, -1), (2, -1), (1, -2), (2, -2)])
def test_conv3d_variants(nco, ks):
    """Every (output tiles, K split) instantiation against the fp32 reference."""
    t = _torch()
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(21)
    shape, cin, cout = (1, 5, 6, 7), 48, 128
    x = t.randn(shape + (cin,), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((27, cin, cout), generator=g) * 0.1).cuda()
    b = (t.randn(cout, generator=g) * 0.1).cuda()
    print(f"[case] conv3d variant nco={nco} ks={ks}", flush=True)
    _, y32 = S.conv3d(x, S.prep_weights(w, 0), b, cout, 3, True, want_bf16=False, want_f32=True,
                      nco=nco, ksplit=ks)
    ref = _ref_conv(x, w, b, 3, True, S.SLOPE)
    err = float((y32.cpu() - ref).abs().max() / ref.abs().max())
    assert err < 1e-5, err


@pytest.mark.parametrize("shape, cin, cout, rows", [
    ((1, 9, 13, 40), 16, 16, 8),      # partial z, y and x bricks
    ((2, 6, 8, 33), 48, 16, 4),       # batch of 2, three 16-channel slices
    ((1, 7, 9, 20), 32, 9, 8),        # fewer than 16 outputs
])
def test_conv3d_c16(shape, cin, cout, rows):
    """The 16-output-channel brick kernel (16x16x32 MFMAs, two taps per K
    step) against the fp32 reference."""
    t = _torch()
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(31)
    x = t.randn(shape + (cin,), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((27, cin, cout), generator=g) * 0.1).cuda()
    b = (t.randn(cout, generator=g) * 0.1).cuda()
    print(f"[case] conv3d c16 {shape} {cin}->{cout} rows={rows}", flush=True)
    _, y32 = S.conv3d(x, S.prep_weights(w, 0), b, cout, 3, True, want_bf16=False, want_f32=True,
                      nco=rows, ksplit=-3)
    ref = _ref_conv(x, w, b, 3, True, S.SLOPE)
    err = float((y32.cpu() - ref).abs().max() / ref.abs().max())
    print(f"[parity] conv3d c16 rel err {err:.2e}", flush=True)
    assert err < 1e-5, err
    # fused [upsample(A), B] read in place
    xa = t.randn((1, 4, 5, 10, 32), generator=g).to(t.bfloat16).cuda()
    xb = t.randn((1, 8, 10, 20, 16), generator=g).to(t.bfloat16).cuda()
    w2 = (t.randn((27, 48, 16), generator=g) * 0.1).cuda()
    cat = t.cat([S._up(xa), xb], dim=-1).contiguous()
    _, y2 = S.conv3d(xa, S.prep_weights(w2, 0), b[:16] if cout == 16 else None, 16, 3, False,
                     want_bf16=False, want_f32=True, xb=xb, up=True, nco=rows, ksplit=-3)
    ref2 = _ref_conv(cat, w2, b[:16] if cout == 16 else None, 3, False, S.SLOPE)
    assert float((y2.cpu() - ref2).abs().max() / ref2.abs().max()) < 1e-5


@pytest.mark.parametrize("cin, cout", [(16, 16), (48, 11)])
def test_conv1_c16(cin, cout):
    """The 16-output-channel 1x1 kernel against the fp32 reference."""
    t = _torch()
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(37)
    shape = (1, 5, 7, 9)                 # 315 voxels: a partial last wave
    x = t.randn(shape + (cin,), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((1, cin, cout), generator=g) * 0.1).cuda()
    b = (t.randn(cout, generator=g) * 0.1).cuda()
    print(f"[case] conv1 c16 {cin}->{cout}", flush=True)
    _, y32 = S.conv3d(x, S.prep_weights(w, 0), b, cout, 1, False, want_bf16=False, want_f32=True,
                      nco=1, ksplit=-3)
    ref = _ref_conv(x, w, b, 1, False, S.SLOPE)
    err = float((y32.cpu() - ref).abs().max() / ref.abs().max())
    assert err < 1e-5, err


def test_conv3d_fused_upsample_concat():
    """Reading [upsample(A), B] in place equals the materialised concatenation
    bit for bit (same kernel, same K order)."""
    t = _torch()
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(5)
    xa = t.randn((2, 3, 4, 5, 64), generator=g).to(t.bfloat16).cuda()
    xb = t.randn((2, 6, 8, 10, 32), generator=g).to(t.bfloat16).cuda()
    w = (t.randn((27, 96, 32), generator=g) * 0.1).cuda()
    b = (t.randn(32, generator=g) * 0.1).cuda()
    wf = S.prep_weights(w, 0)
    cat = t.cat([S._up(xa), xb], dim=-1).contiguous()
    y0, _ = S.conv3d(cat, wf, b, 32, 3, True)
    y1, _ = S.conv3d(xa, wf, b, 32, 3, True, xb=xb, up=True)
    assert t.equal(y0, y1)
    # large enough for the LDS brick kernel
    xa = t.randn((1, 10, 15, 30, 32), generator=g).to(t.bfloat16).cuda()
    xb = t.randn((1, 20, 30, 60, 16), generator=g).to(t.bfloat16).cuda()
    w2 = (t.randn((27, 48, 16), generator=g) * 0.1).cuda()
    wf2 = S.prep_weights(w2, 0)
    cat = t.cat([S._up(xa), xb], dim=-1).contiguous()
    y4, _ = S.conv3d(cat, wf2, b[:16], 16, 3, True)
    y5, _ = S.conv3d(xa, wf2, b[:16], 16, 3, True, xb=xb, up=True)
    assert t.equal(y4, y5)
    _, y6 = S.conv3d(xa, wf2, b[:16], 16, 3, True, want_bf16=False, want_f32=True, xb=xb, up=True)
    ref = _ref_conv(cat, w2, b[:16], 3, True, S.SLOPE)
    assert float((y6.cpu() - ref).abs().max() / ref.abs().max()) < 1e-5
    # upsampled source alone
    y2, _ = S.conv3d(S._up(xa).contiguous(), S.prep_weights(w[:, :64], 0), b, 32, 3, False)
    y3, _ = S.conv3d(xa, S.prep_weights(w[:, :64], 0), b, 32, 3, False, up=True)
    assert t.equal(y2, y3)


def test_pool2_matches_torch():
    t = _torch()
    import torch.nn.functional as F
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(8)
    x = t.randn((2, 6, 8, 4, 48), generator=g).to(t.bfloat16).cuda()
    y = S.pool2(x)
    ref = F.avg_pool3d(x.float().cpu().permute(0, 4, 1, 2, 3), 2).permute(0, 2, 3, 4, 1)
    assert float((y.float().cpu() - ref).abs().max()) <= 2 ** -8 * float(ref.abs().max())


def test_infer_equals_training_forward():
    t = _torch()
    from dosecuda_amd import surrogate as S
    g = t.Generator().manual_seed(3)
    net = S.SurrogateNet().cuda()
    x = t.randn((1, 16, 16, 24, S.IN_CH), generator=g).to(t.bfloat16).cuda()
    with t.no_grad():
        y0 = net(x)
    y1 = net.prepare().infer(x)
    assert t.equal(y0, y1)


def _numpy_gamma(ev, ref, s, dta, dd, cut, rad):
    out = np.full(ref.shape, -1.0)
    nz, ny, nx = ref.shape
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if ref[k, j, i] < cut:
                    continue
                best = np.inf
                for dk in range(-rad, rad + 1):
                    for dj in range(-rad, rad + 1):
                        for di in range(-rad, rad + 1):
                            kk, jj, ii = k + dk, j + dj, i + di
                            if 0 <= kk < nz and 0 <= jj < ny and 0 <= ii < nx:
                                d2 = (di * di + dj * dj + dk * dk) * s * s / dta ** 2
                                dv = (ref[kk, jj, ii] - ev[k, j, i]) ** 2 / dd ** 2
                                best = min(best, d2 + dv)
                out[k, j, i] = math.sqrt(best)
    return out


def test_gamma_known_answers():
    t = _torch()
    from dosecuda_amd import surrogate as S
    s = 2.5
    n = 24
    z, y, x = np.meshgrid(*[np.arange(n) * s] * 3, indexing="ij")
    ref = (np.exp(-((x - 30) ** 2 + (y - 30) ** 2 + (z - 30) ** 2) / 400.0) * 2.0).astype(np.float32)
    R = t.from_numpy(ref).cuda()
    assert S.gamma(R, R, s)[0] == 1.0
    assert S.gamma(R * 1.01, R, s)[0] == 1.0                     # 1% < 2% everywhere
    step = np.where(x < 30.0, 1.0, 0.2).astype(np.float32)
    shifted = np.where(x < 33.75, 1.0, 0.2).astype(np.float32)   # 3.75 mm shift of the edge
    rate, gm = S.gamma(t.from_numpy(shifted).cuda(), t.from_numpy(step).cuda(), s)
    gm = gm.cpu().numpy()
    assert rate < 1.0 and np.all(gm[:, :, :10][gm[:, :, :10] >= 0] <= 1.0)
    # brute force on a small grid
    rng = np.random.default_rng(3)
    a = rng.random((6, 7, 8)).astype(np.float32)
    b = (a + rng.normal(0, 0.02, a.shape)).astype(np.float32)
    rate, gm = S.gamma(t.from_numpy(b).cuda(), t.from_numpy(a).cuda(), 2.0, dta_mm=2.0, dd=0.03,
                       cutoff=0.1, radius_vox=2)
    ref_g = _numpy_gamma(b, a, 2.0, 2.0, 0.03 * a.max(), 0.1 * a.max(), 2)
    np.testing.assert_allclose(gm.cpu().numpy(), ref_g, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("every", [8, 1])
def test_ccc_list_matches_full(photon_model, every):
    """The cone kernel on a voxel list equals the full-volume kernel at those
    voxels bit for bit: a short list (every eighth active voxel, at most 1/6
    of the volume: one item per zenith row and azimuth) and a long one (every
    active voxel, more than 1/6 of the volume: one item per azimuth)."""
    t = _torch()
    from dosecuda_amd import engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    n, s = 40, 2.5
    rho = t.full((n, n, n), 1.00653, device="cuda")
    rho[10:20, 12:30, 8:22] = 0.3
    iso = np.full(3, (n - 1) / 2 * s, np.float32)
    mlc = square_field_mlc(pm, 6.0, 6.0)
    terma, el = t.empty_like(rho), t.empty_like(rho)
    full = engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 40.0, 0.0, 0.0, pm, mlc,
                              terma_out=terma, electron_out=el)
    idx = t.nonzero((~(terma <= 0.01)).view(-1)).view(-1)[::every].to(t.int32).contiguous()
    print(f"[case] ccc list of {idx.numel()} of {rho.numel()} voxels", flush=True)
    assert (idx.numel() <= rho.numel() // 6) == (every == 8)
    out = t.full_like(rho, -7.0)
    S._ccc_list(terma, rho, el, s, np.zeros(3), iso, 100.0, 40.0, 0.0, 0.0, pm, idx, out)
    assert t.equal(out.view(-1)[idx.long()], full.view(-1)[idx.long()])
    mask = t.ones(rho.numel(), dtype=t.bool, device="cuda")
    mask[idx.long()] = False
    assert bool((out.view(-1)[mask] == -7.0).all())


def test_surrogate_trains_and_falls_back(photon_model):
    t = _torch()
    from dosecuda_amd import engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    print("[case] surrogate small train", flush=True)
    net = S.SurrogateNet().cuda()
    # the chunk train() generates first (seed * 1000 + 0)
    X, ch, Y = S.generate(pm, 8, n=64, spacing=(2.0, 3.0), seed=5 * 1000)
    l0 = S.eval_loss(net, X, ch, Y)
    S.train(net, pm, n_samples=8, chunk=8, passes=10, batch=2, n=64, lr=2e-3, spacing=(2.0, 3.0), seed=5)
    l1 = S.eval_loss(net, X, ch, Y)
    print(f"[parity] surrogate training-set loss {l0:.3e} -> {l1:.3e}", flush=True)
    assert l1 < 0.5 * l0
    n, s = 64, 2.5
    rng = np.random.default_rng(9)
    rho = S.random_phantom(n, s, rng)
    iso = np.full(3, (n - 1) / 2 * s, np.float32)
    mlc = square_field_mlc(pm, 8.0, 8.0)
    terma, el = t.empty_like(rho), t.empty_like(rho)
    plain = engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 30.0, 0.0, 0.0, pm, mlc,
                               terma_out=terma, electron_out=el)
    # the target inverts the plain output formula exactly where the dose is positive
    sup = S.superposition_target(plain, el, terma, pm, 100.0)
    back = t.where(sup > 0, (sup + el) * S.dose_scale(pm, 100.0), t.zeros_like(sup))
    act = ~(terma <= 0.01)
    assert float((back - plain).abs().max()) <= 1e-5 * float(plain.max())
    # space-to-depth and depth-to-space are inverse
    assert t.equal(S._d2s(S._s2d(plain)), plain)
    sur = S.SurrogateCCCS(net, pm)
    tm = {}
    # tol < 0: every check fails, every cell is recomputed, the output is the plain dose
    d = sur.dose(terma, rho, el, s, np.zeros(3), iso, 100.0, 30.0, 0.0, 0.0, mlc, stride=4,
                 tol=-1.0, timings=tm)
    print(f"[parity] surrogate fallback-all: {tm}", flush=True)
    assert t.equal(d[act], plain[act])
    rate, _ = S.gamma(d, plain, s)
    assert rate == 1.0


def test_surrogate_weights_roundtrip(tmp_path):
    t = _torch()
    from dosecuda_amd import surrogate as S
    net = S.SurrogateNet().cuda()
    net.out_scale.fill_(3.5)
    p = str(tmp_path / "w.safetensors")
    S.save(net, p)
    net2 = S.load(p)
    for (k, a), (_, b) in zip(net.state_dict().items(), net2.state_dict().items()):
        assert t.equal(a.cpu(), b.cpu()), k


def test_pack_and_finish_match_torch(photon_model):
    """The HIP input packing equals the torch restatement (make_input) bit for
    bit, on a volume whose sizes are not multiples of 16; the HIP finish
    equals the torch formula of the plain CCCS's output."""
    t = _torch()
    from dosecuda_amd import engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    shape, s = (37, 45, 50), 2.5
    rho = t.full(shape, 1.00653, device="cuda")
    rho[5:20, 10:30, 20:40] = 0.35
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    terma, el = t.empty_like(rho), t.empty_like(rho)
    engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 70.0, 0.0, 0.0, pm, square_field_mlc(pm, 6.0, 6.0),
                       terma_out=terma, electron_out=el)
    net = S.SurrogateNet().cuda()
    net.out_scale.fill_(2.5)
    sur = S.SurrogateCCCS(net, pm)
    d = sur.predict(terma, rho, el, 70.0, 0.0, 100.0, s)
    b = sur._buffers(shape, terma.device, sur.stride)
    x_ref, scale = S.make_input(terma, rho, S.beam_direction(70.0, 0.0), s)
    assert t.equal(b["x"][0], x_ref)
    assert t.equal(b["tmax"][0], scale)
    y = net(x_ref[None])[0, ..., :S.N_SUB]
    sup = S._d2s(y)[:shape[0], :shape[1], :shape[2]] * (scale * net.out_scale)
    dd = (sup + el) * S.dose_scale(pm, 100.0)
    ref = t.where((terma > 0.01) & (dd > 0), dd, t.zeros_like(dd))
    assert t.equal(d, ref)
    assert float(b["dmax"][0]) == float(ref.max())


def _torch_checked(pred, plain, terma, stride, tol, dilate=2):
    """The lattice check and fallback restated with torch operations: each
    cell of stride^3 voxels is checked at its low corner when that voxel is
    active, else at its centre voxel (offset stride // 2) when that is in the
    volume and active, else it has no check voxel and is recomputed."""
    import torch
    import torch.nn.functional as F
    active = ~(terma <= 0.01)
    nz, ny, nx = active.shape
    L = [-(-n // stride) for n in active.shape]
    h = stride // 2
    cells = lambda v: v[::stride, ::stride, ::stride]
    corner = cells(active)
    centre = torch.zeros(L, dtype=torch.bool, device=active.device)
    if h > 0:
        ca = active[h::stride, h::stride, h::stride]
        centre[:ca.shape[0], :ca.shape[1], :ca.shape[2]] = ca
    centre &= ~corner
    err = lambda v: cells(v)
    dmax = torch.clamp(pred.max(), min=1e-12)
    diff = (pred - plain).abs() > tol * dmax
    dc = torch.zeros(L, dtype=torch.bool, device=active.device)
    if h > 0:
        e = diff[h::stride, h::stride, h::stride]
        dc[:e.shape[0], :e.shape[1], :e.shape[2]] = e
    bad = (corner & err(diff)) | (centre & dc)
    checked = corner | centre
    pb = bad.float()[None, None]
    if dilate == 2:
        flag = F.max_pool3d(pb, kernel_size=3, stride=1, padding=1)
    elif dilate == 1:
        k = torch.zeros((1, 1, 3, 3, 3), device=pb.device)
        k[0, 0, 1, 1, :] = 1
        k[0, 0, 1, :, 1] = 1
        k[0, 0, :, 1, 1] = 1
        flag = (F.conv3d(pb, k, padding=1) > 0).float()
    else:
        flag = pb
    pool = lambda v: F.max_pool3d(v.float()[None, None], kernel_size=stride, stride=stride, ceil_mode=True)
    flag = torch.maximum(flag, pool(active) * (~checked).float()[None, None])
    up = flag.repeat_interleave(stride, 2).repeat_interleave(stride, 3).repeat_interleave(stride, 4)
    fb = (up[0, 0, :nz, :ny, :nx] > 0) & active
    return (torch.where(fb, plain, pred), int(checked.sum()), int(fb.sum()), int(centre.sum()),
            int((centre & dc).sum()))


@pytest.mark.parametrize("stride, dilate, field", [(3, 2, 7.0), (4, 2, 7.0), (4, 1, 7.0), (4, 0, 7.0), (4, 1, 3.0),
                                                   (3, 2, 3.0)])
def test_device_check_matches_torch(photon_model, stride, dilate, field):
    """The device check equals its torch restatement: lattice and fallback
    counts and the checked dose bit for bit.  In a 7 cm field every voxel
    has TERMA > 0.01; in a 3 cm field cells across the field edge have an
    inactive corner and are checked at their centre voxel, and those of the
    upper half are made to fail there."""
    t = _torch()
    from dosecuda_amd import engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    shape, s = (42, 40, 44), 2.5
    rho = t.full(shape, 1.00653, device="cuda")
    rho[8:20, 5:25, 10:30] = 0.3
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    terma, el = t.empty_like(rho), t.empty_like(rho)
    plain = engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 200.0, 0.0, 0.0, pm,
                               square_field_mlc(pm, field, field), terma_out=terma, electron_out=el)
    pred = plain.clone()
    pred[10:18, 10:20, 12:22] *= 1.05          # fails a 1% check where the dose is high
    pred[25:35, 25:35, 25:35] *= 1.002         # passes
    pred[30:32, 2:6, 2:6] = 0.0                # isolated failures
    # cells of the upper half whose corner voxel is inactive (across the
    # field edge, checked at their centre voxel) 2% of the maximum high
    act = ~(terma <= 0.01)
    edge = (~act[::stride, ::stride, ::stride]).repeat_interleave(stride, 0).repeat_interleave(stride, 1)
    edge = edge.repeat_interleave(stride, 2)[:shape[0], :shape[1], :shape[2]] & act
    edge[21:] = False
    pred[edge] += 0.02 * float(plain.max())
    want, n_lat, n_fb, n_centre, n_centre_bad = _torch_checked(pred, plain, terma, stride, 0.01, dilate)
    if field < 5.0:
        assert n_centre > n_centre_bad > 0
    else:
        assert n_centre == 0
    sur = S.SurrogateCCCS(S.SurrogateNet().cuda(), pm, stride=stride, dilate=dilate)
    b = sur._buffers(shape, terma.device, stride)
    b["dmax"].fill_(float(pred.max()))
    got = pred.clone()
    keep: list = []
    m = engine.photon_args(pm, keep)
    from dosecuda_amd import _lib
    _lib.check(_lib.load().dh_sur_check_dev(0, _lib.grid(shape, s, np.zeros(3)),
                                            _lib.photon_cp(iso, 100.0, 200.0, 0.0, 0.0), m, _lib.ptr(terma),
                                            _lib.ptr(rho), _lib.ptr(el), _lib.ptr(got), _lib.ptr(b["dmax"]),
                                            stride, 0.01, dilate, _lib.ptr(b["work"]), b["work"].numel(),
                                            _lib.ptr(b["plain"]), S._st()))
    lat, fb = sur.counts(shape, stride=stride)
    print(f"[parity] device check stride {stride} dilate {dilate}: lattice {lat} fallback {fb} "
          f"(torch {n_lat}, {n_fb}, {n_centre} at cell centres)", flush=True)
    assert (lat, fb) == (n_lat, n_fb)
    assert 0 < fb < int((~(terma <= 0.01)).sum())
    assert t.equal(got, want)


def test_native_dose_equals_torch_path(photon_model):
    """One library call (dh_sur_dose_dev: native forward) equals the
    Python-driven network + device check bit for bit, and the photon path
    with the surrogate equals plain TERMA + dh_sur_dose_dev."""
    t = _torch()
    from dosecuda_amd import _lib, engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    shape, s = (40, 50, 36), 2.5
    rho = t.full(shape, 1.00653, device="cuda")
    rho[10:25, 5:30, 8:20] = 0.3
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    mlc = square_field_mlc(pm, 6.0, 6.0)
    terma, el = t.empty_like(rho), t.empty_like(rho)
    engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 120.0, 0.0, 0.0, pm, mlc, terma_out=terma,
                       electron_out=el)
    g = t.Generator().manual_seed(2)
    net = S.SurrogateNet().cuda()
    with t.no_grad():
        for p in net.parameters():
            p.mul_(0.5).add_(0.01 * t.randn(p.shape, generator=g).cuda())
    net.out_scale.fill_(1.7)
    sur = S.SurrogateCCCS(net, pm, stride=4, tol=0.02)
    a = sur.dose(terma, rho, el, s, np.zeros(3), iso, 100.0, 120.0, 0.0, 0.0)
    b = sur.dose_torch(terma, rho, el, s, np.zeros(3), iso, 100.0, 120.0, 0.0, 0.0)
    assert t.equal(a, b)
    # native forward alone
    x = t.randn((1, 16, 24, 16, S.IN_CH), generator=g).to(t.bfloat16).cuda()
    y = t.empty((1, 16, 24, 16, 16), device="cuda")
    _lib.check(_lib.load().dh_sur_net_forward_dev(sur.native.handle, _lib.ptr(x), 16, 24, 16, _lib.ptr(y), S._st()))
    assert t.equal(y, net.infer(x))
    # whole photon path with the surrogate
    d = t.empty_like(rho)
    keep: list = []
    m = engine.photon_args(pm, keep)
    rows = np.ascontiguousarray(np.asarray(mlc, dtype=np.float32)[:, :4])
    _lib.check(_lib.load().dh_photon_dose_surrogate_dev(
        0, sur.native.handle, _lib.grid(shape, s, np.zeros(3)), _lib.photon_cp(iso, 100.0, 120.0, 0.0, 0.0), m,
        _lib.ptr(rows), rows.shape[0], _lib.ptr(rho), _lib.ptr(d), None, None, 4, 0.02, sur.dilate, S._st()))
    # (its depth comes from the beam's-eye-view grid: the same TERMA as photon_terma's at step 0)
    tf, ef = engine.photon_terma(rho, s, np.zeros(3), iso, 100.0, 120.0, 0.0, 0.0, pm, mlc, step=0.0)
    a2 = sur.dose(tf, rho, ef, s, np.zeros(3), iso, 100.0, 120.0, 0.0, 0.0)
    assert t.equal(d, a2)
    # at step 1 mm the TERMA is the plain path's bit for bit
    t1, e1 = engine.photon_terma(rho, s, np.zeros(3), iso, 100.0, 120.0, 0.0, 0.0, pm, mlc, step=1.0)
    assert t.equal(t1, terma) and t.equal(e1, el)


@pytest.mark.parametrize("shape, n_cp", [((40, 50, 36), 3), ((96, 96, 96), 9)])
def test_dose_batch_equals_single(photon_model, shape, n_cp):
    """dh_sur_dose_batch_dev (control points sharing network launches, up
    to 8 per batch) gives each point the dose of dh_sur_dose_dev on that
    point alone, bit for bit; the larger grid takes the LDS-brick and
    16-channel convolution kernels at full resolution, and 9 points take
    two batches."""
    t = _torch()
    from dosecuda_amd import engine, surrogate as S
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    s = 2.5
    rho = t.full(shape, 1.00653, device="cuda")
    rho[shape[0] // 4:shape[0] // 2, shape[1] // 8:shape[1] // 2, shape[2] // 4:shape[2] // 2] = 0.3
    iso = np.array([(shape[2] - 1) / 2 * s, (shape[1] - 1) / 2 * s, (shape[0] - 1) / 2 * s], np.float32)
    mlc = square_field_mlc(pm, 6.0, 6.0)
    sur = S.SurrogateCCCS(S.load(S.WEIGHTS), pm, stride=3, tol=0.01)
    cps, termas, els = [], [], []
    for i in range(n_cp):
        ga, ca = 360.0 * i / n_cp, 10.0 * (i % 3)
        terma, el = t.empty_like(rho), t.empty_like(rho)
        engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, ga, ca, 0.0, pm, mlc, terma_out=terma,
                           electron_out=el)
        cps.append((iso, 100.0 + i, ga, ca, 0.0))
        termas.append(terma)
        els.append(el)
    print(f"[case] surrogate dose batch shape={shape} points={n_cp}", flush=True)
    got = sur.dose_batch(termas, rho, els, s, np.zeros(3), cps)
    for i, cp in enumerate(cps):
        want = sur.dose(termas[i], rho, els[i], s, np.zeros(3), *cp)
        assert float(want.max()) > 0.0
        assert t.equal(got[i], want), f"control point {i}"
    with pytest.raises(ValueError):
        sur.dose_batch(termas[:1], rho, els, s, np.zeros(3), cps)
    with pytest.raises(ValueError):
        sur.dose_batch(termas[:2], rho, els[:2], s, np.zeros(3), cps[:2], outs=[got[0], got[0]])
    # an output over an input volume is refused, by the wrapper and by the C ABI
    with pytest.raises(ValueError, match="overlap"):
        sur.dose_batch(termas[:2], rho, els[:2], s, np.zeros(3), cps[:2], outs=[got[0], termas[1]])
    import ctypes as C
    from dosecuda_amd import _lib
    keep: list = []
    m = engine.photon_args(pm, keep)
    arr = (_lib.PhotonCP * 2)(*[_lib.photon_cp(*cp) for cp in cps[:2]])
    P = lambda vs: (C.c_void_p * 2)(*[v.data_ptr() for v in vs])
    for outs_bad in ([got[0], els[0]], [got[0], rho], [got[1], got[1]]):
        rc = _lib.load().dh_sur_dose_batch_dev(sur.native.handle, _lib.grid(rho.shape, s, np.zeros(3)), 2, arr, m,
                                               P(termas[:2]), _lib.ptr(rho), P(els[:2]), 3, 0.01, 1, P(outs_bad),
                                               C.c_void_p(t.cuda.current_stream().cuda_stream))
        assert rc == _lib.DH_E_INVALID and "overlap" in _lib.last_error()
    # the batch (with its side stream) captured as one HIP graph; replays
    # interleaved with eager work must reproduce the eager doses
    outs = [t.empty_like(rho) for _ in cps]
    sur.dose_batch(termas, rho, els, s, np.zeros(3), cps, outs=outs)
    t.cuda.synchronize()
    gr = t.cuda.CUDAGraph()
    with t.cuda.graph(gr):
        sur.dose_batch(termas, rho, els, s, np.zeros(3), cps, outs=outs)
    for r in range(3):
        for o in outs:
            o.fill_(-1.0)
        gr.replay()
        t.cuda.synchronize()
        for i in range(n_cp):
            assert t.equal(outs[i], got[i]), f"replay {r}, control point {i}"
        sur.dose(termas[0], rho, els[0], s, np.zeros(3), *cps[0])


@pytest.mark.parametrize("ga, ca, ta", [(0.0, 0.0, 0.0), (33.0, 15.0, 0.0), (250.0, 0.0, 20.0)])
def test_bev_depth_terma_close_to_raytrace(photon_model, ga, ca, ta):
    """The checked surrogate path's radiological depth (fan-line integrals on
    a beam's-eye-view grid: photon_terma at step 0) gives a TERMA close to the
    reference ray trace's (step 1 mm): on water within 1% of the local value
    wherever TERMA exceeds 1% of its maximum two voxels or more inside the
    volume's faces, and within 1.5% on the faces too wherever it exceeds 10%
    of its maximum (lattice lines that miss the volume's voxel centres at a
    face voxel's distance are left out of its interpolation, so face voxels
    are extrapolated from the lines inside; depths below the electron term's
    reach come from the 1 mm ray trace); on a 72^3 lung/bone
    phantom within 1.5% of the maximum everywhere and 0.1% on average;
    replaying the step-0 TERMA from a HIP graph gives the eager result."""
    t = _torch()
    from dosecuda_amd import engine, phantoms
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    mlc = square_field_mlc(pm, 8.0, 8.0)
    print(f"[case] bev depth gantry={ga} collimator={ca} couch={ta}", flush=True)
    case = phantoms.imrt_lung_bone_case(seed=1002, n=72, spacing=2.5)
    lung = t.from_numpy(phantoms.to_density(case)).cuda()
    water = t.full_like(lung, 1.0)
    for name, rho in (("water", water), ("lung/bone", lung)):
        args = (rho, case.spacing, case.origin, case.iso, 100.0, ga, ca, ta, pm, mlc)
        ref, _ = engine.photon_terma(*args, step=1.0)
        got, el = engine.photon_terma(*args, step=0.0)
        tmax = float(ref.max())
        m = ref > 0.01 * tmax
        dev_abs = (got - ref).abs()
        inner = t.zeros_like(m)
        inner[2:-2, 2:-2, 2:-2] = True
        mi = m & inner
        rel = float((dev_abs[mi] / ref[mi]).max())
        m10 = ref > 0.1 * tmax
        rel_faces = float((dev_abs[m10] / ref[m10]).max())
        mx, mean = float(dev_abs.max()) / tmax, float(dev_abs[m].mean()) / tmax
        print(f"[parity] bev depth {name}: max rel {rel:.2e} (with the faces, above 10% of max {rel_faces:.2e}), "
              f"max {mx:.2e} / mean {mean:.2e} of max TERMA", flush=True)
        if name == "water":
            assert rel <= 0.01 and rel_faces <= 0.015
        else:
            assert mx <= 0.015 and mean <= 1e-3
    eager = got.clone()
    g = t.cuda.CUDAGraph()
    out_t, out_e = t.empty_like(got), t.empty_like(el)
    engine.photon_terma(*args, step=0.0, terma_out=out_t, electron_out=out_e)
    t.cuda.synchronize()
    with t.cuda.graph(g):
        engine.photon_terma(*args, step=0.0, terma_out=out_t, electron_out=out_e)
    for _ in range(3):
        out_t.zero_()
        g.replay()
        t.cuda.synchronize()
        assert t.equal(out_t, eager)


@pytest.mark.parametrize("step", [0.0, 2.5])
def test_terma_batch_equals_single(photon_model, step):
    """photon_terma_batch (density padded once, points alternating over two
    streams with their own depth buffers) gives every point's TERMA and
    electron volumes bit for bit as photon_terma of that point alone, eager
    and from a HIP graph replayed between eager work; outputs that overlap
    each other or the density are refused."""
    t = _torch()
    from dosecuda_amd import engine, phantoms, _lib
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    case = phantoms.imrt_lung_bone_case(seed=1002, n=56, spacing=2.5)
    rho = t.from_numpy(phantoms.to_density(case)).cuda()
    mlcs = [square_field_mlc(pm, 8.0, 6.0), square_field_mlc(pm, 4.0, 10.0), square_field_mlc(pm, 10.0, 10.0)]
    cps = [(case.iso, 100.0, 0.0, 0.0, 0.0), (case.iso, 80.0, 110.0, 15.0, 0.0), (case.iso, 50.0, 250.0, 0.0, 20.0)]
    print(f"[case] terma batch step={step}", flush=True)
    ref = [engine.photon_terma(rho, case.spacing, case.origin, *cp, pm, m, step=step) for cp, m in zip(cps, mlcs)]
    ts, es = engine.photon_terma_batch(rho, case.spacing, case.origin, cps, pm, mlcs, step=step)
    for i, (rt, re) in enumerate(ref):
        assert t.equal(ts[i], rt) and t.equal(es[i], re), f"point {i}"
    g = t.cuda.CUDAGraph()
    with t.cuda.graph(g):
        engine.photon_terma_batch(rho, case.spacing, case.origin, cps, pm, mlcs, step=step, terma_outs=ts,
                                  electron_outs=es)
    for r in range(3):
        for v in (*ts, *es):
            v.fill_(-1.0)
        g.replay()
        t.cuda.synchronize()
        for i, (rt, re) in enumerate(ref):
            assert t.equal(ts[i], rt) and t.equal(es[i], re), f"replay {r}, point {i}"
        engine.photon_terma(rho, case.spacing, case.origin, *cps[0], pm, mlcs[0], step=step)
    for bad in ((ts[0], ts[0]), (rho, es[1])):
        with pytest.raises(ValueError, match="overlap"):
            engine.photon_terma_batch(rho, case.spacing, case.origin, cps[:2], pm, mlcs[:2], step=step,
                                      terma_outs=[ts[0], bad[0]], electron_outs=[es[0], bad[1]])
