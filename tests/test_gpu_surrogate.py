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


def test_ccc_list_matches_full(photon_model):
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
    idx = t.nonzero((~(terma <= 0.01)).view(-1)).view(-1)[::3].to(t.int32).contiguous()
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
    X, Y = S.generate(pm, 6, n=64, spacing=2.5, seed=5)
    net = S.SurrogateNet().cuda()
    hist = S.train(net, X, Y, steps=40, batch=4, crop=16, lr=2e-3)
    print(f"[parity] surrogate loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    assert hist[-1] < hist[0]
    n, s = 64, 2.5
    rng = np.random.default_rng(9)
    rho = S.random_phantom(n, s, rng)
    iso = np.full(3, (n - 1) / 2 * s, np.float32)
    mlc = square_field_mlc(pm, 8.0, 8.0)
    terma, el = t.empty_like(rho), t.empty_like(rho)
    plain = engine.photon_dose(rho, s, np.zeros(3), iso, 100.0, 30.0, 0.0, 0.0, pm, mlc,
                               terma_out=terma, electron_out=el)
    sur = S.SurrogateCCCS(net, pm)
    tm = {}
    # tol < 0: every check fails, every cell is recomputed, the output is the plain dose
    d = sur.dose(terma, rho, el, s, np.zeros(3), iso, 100.0, 30.0, 0.0, 0.0, mlc, stride=4,
                 tol=-1.0, timings=tm)
    act = ~(terma <= 0.01)
    print(f"[parity] surrogate fallback-all: {tm}", flush=True)
    assert t.equal(d[act], plain[act])
    rate, _ = S.gamma(d, plain, s)
    assert rate == 1.0
