"""Learned CCCS surrogate with plain-CCCS fallback.

A small 3-D convolutional encoder-decoder (U-Net, widths 16/32/64/128, about
1.5 M parameters) predicts the collapsed-cone dose of one photon beam from
its TERMA, density and electron-contamination volumes.  The network runs on
a 2x average-pooled grid (a 128^3 volume becomes a 64^3 patch) and its output
is trilinearly upsampled.  Every convolution (forward, input gradient and
weight gradient) is a hand-written BF16 MFMA kernel of libdosehip.so with
FP32 accumulation; torch supplies memory, pooling/upsampling/concatenation
glue and the Adam update of the FP32 master weights.

Trust, on the timed path: the plain CCCS (`k_ccc`) is evaluated on a sparse
lattice of voxels (every `stride`-th voxel along each axis among voxels with
TERMA > 0.01); every lattice cell whose surrogate value differs from the
plain value by more than `tol` of the maximum, together with its neighbour
cells, is recomputed with the plain CCCS voxel by voxel.  The full plain
CCCS and the 2%/2 mm gamma index are the judge, off the timed path.

No reference counterpart: DoseCUDA computes only the plain CCCS
(IMRTClasses.cu:252-350).
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib, engine

SLOPE = 0.1          # LeakyReLU negative slope
IN_CH = 16           # input channels (T, rho, E + zero padding to a multiple of 16)


def _st(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ceil(n, m):
    return (n + m - 1) // m * m


# ---------------------------------------------------------------------------
# kernel wrappers
# ---------------------------------------------------------------------------
def prep_weights(w32: torch.Tensor, mode: int) -> torch.Tensor:
    """fp32 [taps, cin, cout] -> bf16 [taps, cout_pad32, cin] (mode 0) or
    tap-reversed [taps, cin_pad32, cout] (mode 1)."""
    taps, cin, cout = w32.shape
    rows = _ceil(cout if mode == 0 else cin, 32)
    out = torch.empty((taps, rows, cin if mode == 0 else cout), dtype=torch.bfloat16, device=w32.device)
    _lib.check(_lib.load().dh_nn_prep_weights_dev(_lib.ptr(w32), taps, cin, cout, mode, rows,
                                                  _lib.ptr(out), _st()))
    return out


def conv3d(x: torch.Tensor, wprep: torch.Tensor, bias, cout: int, ksize: int, act: bool,
           want_bf16=True, want_f32=False):
    n, d, h, w, cin = x.shape
    y = torch.empty((n, d, h, w, cout), dtype=torch.bfloat16, device=x.device) if want_bf16 else None
    y32 = torch.empty((n, d, h, w, cout), dtype=torch.float32, device=x.device) if want_f32 else None
    _lib.check(_lib.load().dh_nn_conv3d_dev(_lib.ptr(x), n, d, h, w, cin, _lib.ptr(wprep),
                                            _lib.ptr(bias), cout, wprep.shape[1], ksize,
                                            1 if act else 0, SLOPE, _lib.ptr(y), _lib.ptr(y32), _st()))
    return y, y32


def wgrad(x: torch.Tensor, dy: torch.Tensor, ksize: int) -> torch.Tensor:
    n, d, h, w, cin = x.shape
    cout = dy.shape[-1]
    dw = torch.zeros((ksize ** 3, cin, cout), dtype=torch.float32, device=x.device)
    _lib.check(_lib.load().dh_nn_conv3d_wgrad_dev(_lib.ptr(x), _lib.ptr(dy), n, d, h, w, cin, cout,
                                                  ksize, _lib.ptr(dw), _st()))
    return dw


class _Conv(torch.autograd.Function):
    """y = act(conv(x, w) + b) on channels-last bf16 x; w fp32 [taps, cin, cout]."""

    @staticmethod
    def forward(ctx, x, w, b, ksize, act, out32):
        x = x.contiguous()
        wf = prep_weights(w, 0)
        y, y32 = conv3d(x, wf, b, w.shape[2], ksize, act, want_bf16=not out32, want_f32=out32)
        out = y32 if out32 else y
        ctx.save_for_backward(x, w, out)
        ctx.ksize, ctx.act = ksize, act
        return out

    @staticmethod
    def backward(ctx, gy):
        x, w, out = ctx.saved_tensors
        g = gy.float()
        if ctx.act:
            g = g * torch.where(out.float() > 0, 1.0, SLOPE)
        gb = g.to(torch.bfloat16).contiguous()
        wd = prep_weights(w, 1)
        dx, _ = conv3d(gb, wd, None, w.shape[1], ctx.ksize, False)
        dw = wgrad(x, gb, ctx.ksize)
        db = g.sum(dim=(0, 1, 2, 3))
        return dx, dw, db, None, None, None


class Conv(torch.nn.Module):
    def __init__(self, cin, cout, ksize=3, act=True, out32=False, gen=None):
        super().__init__()
        fan_in = cin * ksize ** 3
        w = torch.randn((ksize ** 3, cin, cout), generator=gen) * math.sqrt(2.0 / fan_in)
        self.w = torch.nn.Parameter(w)
        self.b = torch.nn.Parameter(torch.zeros(cout))
        self.ksize, self.act, self.out32 = ksize, act, out32

    def forward(self, x):
        return _Conv.apply(x, self.w, self.b, self.ksize, self.act, self.out32)


def _pool(x):
    y = F.avg_pool3d(x.permute(0, 4, 1, 2, 3), 2)
    return y.permute(0, 2, 3, 4, 1).contiguous()


def _up(x):
    y = F.interpolate(x.permute(0, 4, 1, 2, 3), scale_factor=2, mode="nearest")
    return y.permute(0, 2, 3, 4, 1).contiguous()


class SurrogateNet(torch.nn.Module):
    """U-Net on channels-last bf16 tensors [N, D, H, W, 16] (D, H, W
    multiples of 8); output fp32 [N, D, H, W] (normalised dose)."""

    def __init__(self, widths=(16, 32, 64, 128), seed=1004):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        a, b, c, d = widths
        self.e0 = torch.nn.ModuleList([Conv(IN_CH, a, gen=g), Conv(a, a, gen=g)])
        self.e1 = torch.nn.ModuleList([Conv(a, b, gen=g), Conv(b, b, gen=g)])
        self.e2 = torch.nn.ModuleList([Conv(b, c, gen=g), Conv(c, c, gen=g)])
        self.bt = torch.nn.ModuleList([Conv(c, d, gen=g), Conv(d, d, gen=g)])
        self.d2 = torch.nn.ModuleList([Conv(d + c, c, gen=g), Conv(c, c, gen=g)])
        self.d1 = torch.nn.ModuleList([Conv(c + b, b, gen=g), Conv(b, b, gen=g)])
        self.d0 = torch.nn.ModuleList([Conv(b + a, a, gen=g), Conv(a, a, gen=g)])
        # 1x1 head; 16 output channels of which channel 0 is the dose, so the
        # input-gradient convolution keeps a multiple-of-16 channel count
        self.head = Conv(a, 16, ksize=1, act=False, out32=True, gen=g)

    @staticmethod
    def _seq(mods, x):
        for m in mods:
            x = m(x)
        return x

    def forward(self, x):
        s0 = self._seq(self.e0, x)
        s1 = self._seq(self.e1, _pool(s0))
        s2 = self._seq(self.e2, _pool(s1))
        t = self._seq(self.bt, _pool(s2))
        t = self._seq(self.d2, torch.cat([_up(t), s2], dim=-1))
        t = self._seq(self.d1, torch.cat([_up(t), s1], dim=-1))
        t = self._seq(self.d0, torch.cat([_up(t), s0], dim=-1))
        return self.head(t)[..., 0]

    def n_params(self):
        return sum(p.numel() for p in self.parameters())

    def flops(self, vox):
        """Multiply-add FLOPs of one forward pass over `vox` voxels of the
        network's working grid (2 * taps * cin * cout per output voxel)."""
        tot = 0.0
        for name, mods, scale in (("e0", self.e0, 1), ("e1", self.e1, 8), ("e2", self.e2, 64),
                                  ("bt", self.bt, 512), ("d2", self.d2, 64), ("d1", self.d1, 8),
                                  ("d0", self.d0, 1)):
            for m in mods:
                t, ci, co = m.w.shape
                tot += 2.0 * t * ci * co * vox / scale
        t, ci, co = self.head.w.shape
        return tot + 2.0 * t * ci * co * vox


# ---------------------------------------------------------------------------
# inputs / targets
# ---------------------------------------------------------------------------
def make_input(terma: torch.Tensor, density: torch.Tensor, electron: torch.Tensor):
    """Full-resolution fp32 (D, H, W) volumes -> channels-last bf16
    [1, D/2, H/2, W/2, 16] on the 2x pooled grid and the TERMA scale."""
    scale = torch.clamp(terma.max(), min=1e-6)
    v = torch.stack([terma / scale, density, electron / scale], 0)[None]      # [1, 3, D, H, W]
    v = F.avg_pool3d(v, 2)
    x = torch.zeros((1,) + tuple(v.shape[2:]) + (IN_CH,), dtype=torch.bfloat16, device=v.device)
    x[..., :3] = v[0].permute(1, 2, 3, 0).to(torch.bfloat16)
    return x, scale


def dose_scale(model, mu):
    return float(np.float32(model.mu_calibration)) * float(mu)


def random_phantom(n, spacing, rng: np.random.Generator, device="cuda"):
    """Water cube with random lung ellipsoids, bone cylinders along z and
    bone slabs (densities from the default HU table range)."""
    z = torch.arange(n, device=device, dtype=torch.float32)[:, None, None] * spacing
    y = torch.arange(n, device=device, dtype=torch.float32)[None, :, None] * spacing
    x = torch.arange(n, device=device, dtype=torch.float32)[None, None, :] * spacing
    ext = n * spacing
    rho = torch.full((n, n, n), 1.00653, device=device)
    for _ in range(int(rng.integers(0, 3))):
        c = rng.uniform(0.2, 0.8, 3) * ext
        a = rng.uniform(15.0, 60.0, 3)
        m = ((x - c[0]) / a[0]) ** 2 + ((y - c[1]) / a[1]) ** 2 + ((z - c[2]) / a[2]) ** 2 <= 1.0
        rho[m.expand(n, n, n)] = float(rng.uniform(0.2, 0.45))
    for _ in range(int(rng.integers(0, 3))):
        c = rng.uniform(0.2, 0.8, 2) * ext
        r = rng.uniform(8.0, 25.0)
        m = (x - c[0]) ** 2 + (y - c[1]) ** 2 <= r * r
        rho[m.expand(n, n, n)] = float(rng.uniform(1.3, 1.9))
    if rng.random() < 0.3:
        d0 = rng.uniform(0.1, 0.7) * ext
        m = (y >= d0) & (y < d0 + rng.uniform(10.0, 30.0))
        rho[m.expand(n, n, n)] = float(rng.uniform(1.4, 1.75))
    return rho.contiguous()


def generate(model, n_samples, n=128, spacing=2.5, seed=1004, log=None):
    """Training data from the plain CCCS: random phantoms, gantry angles and
    square fields of 3-20 cm; returns inputs [S, n/2, n/2, n/2, 16] bf16 and
    normalised targets [S, n/2, n/2, n/2] bf16 on the pooled grid."""
    from .beam_models import square_field_mlc
    rng = np.random.default_rng(seed)
    half = n // 2
    X = torch.empty((n_samples, half, half, half, IN_CH), dtype=torch.bfloat16, device="cuda")
    Y = torch.empty((n_samples, half, half, half), dtype=torch.bfloat16, device="cuda")
    iso = np.full(3, (n - 1) / 2 * spacing, np.float32)
    terma = torch.empty((n, n, n), device="cuda")
    el = torch.empty_like(terma)
    dose = torch.empty_like(terma)
    t0 = time.time()
    for s in range(n_samples):
        rho = random_phantom(n, spacing, rng)
        ga = float(rng.uniform(0.0, 360.0))
        fs = float(rng.uniform(3.0, 20.0))
        mlc = square_field_mlc(model, fs, fs)
        engine.photon_dose(rho, spacing, np.zeros(3), iso, 100.0, ga, 0.0, 0.0, model, mlc,
                           out=dose, terma_out=terma, electron_out=el)
        x, scale = make_input(terma, rho, el)
        X[s] = x[0]
        tgt = F.avg_pool3d((dose / (dose_scale(model, 100.0) * scale))[None, None], 2)[0, 0]
        Y[s] = tgt.to(torch.bfloat16)
        if log and (s + 1) % max(1, n_samples // 10) == 0:
            log(f"generated {s + 1}/{n_samples} samples ({time.time() - t0:.1f} s)")
    return X, Y


def train(net: SurrogateNet, X, Y, steps=2000, batch=8, crop=32, lr=2e-3, seed=1004, log=None):
    """Adam on the FP32 master weights; random crop^3 patches of the pooled
    training volumes; loss = mean squared error weighted towards dose."""
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=steps, pct_start=0.1)
    g = torch.Generator(device="cpu").manual_seed(seed)
    S, D = X.shape[0], X.shape[1]
    hist = []
    t0 = time.time()
    for it in range(steps):
        idx = torch.randint(0, S, (batch,), generator=g)
        o = torch.randint(0, D - crop + 1, (batch, 3), generator=g)
        xb = torch.stack([X[i, a:a + crop, b:b + crop, c:c + crop] for i, (a, b, c) in zip(idx.tolist(), o.tolist())])
        yb = torch.stack([Y[i, a:a + crop, b:b + crop, c:c + crop] for i, (a, b, c) in zip(idx.tolist(), o.tolist())]).float()
        pred = net(xb.contiguous())
        wgt = 1.0 + 4.0 * (yb > 0.1 * yb.amax(dim=(1, 2, 3), keepdim=True))
        loss = (wgt * (pred - yb) ** 2).mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if (it + 1) % max(1, steps // 20) == 0:
            hist.append(float(loss.item()))
            if log:
                log(f"step {it + 1}/{steps} loss {hist[-1]:.3e} ({time.time() - t0:.1f} s)")
    return hist


# ---------------------------------------------------------------------------
# inference with the plain fallback
# ---------------------------------------------------------------------------
class SurrogateCCCS:
    def __init__(self, net: SurrogateNet, model):
        self.net = net
        self.model = model

    @torch.no_grad()
    def predict(self, terma, density, electron, mu):
        """Surrogate dose (fp32, full resolution) of one beam."""
        D, H, W = terma.shape
        x, scale = make_input(terma, density, electron)
        y = self.net(x)                                        # [1, D/2, H/2, W/2]
        y = F.interpolate(y[None], size=(D, H, W), mode="trilinear", align_corners=False)[0, 0]
        return torch.clamp(y, min=0.0) * (dose_scale(self.model, mu) * scale)

    def dose(self, terma, density, electron, spacing, origin, iso, mu, ga, ca, ta, mlc,
             stride=4, tol=0.01, timings=None):
        """Surrogate dose checked on a lattice against the plain CCCS, with the
        plain CCCS recomputed voxel by voxel wherever the check fails."""
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if timings is not None else None
        if ev:
            ev[0].record()
        d = self.predict(terma, density, electron, mu)
        if ev:
            ev[1].record()
        active = ~(terma <= 0.01)
        lat = torch.zeros_like(active)
        lat[::stride, ::stride, ::stride] = True
        lat_idx = torch.nonzero((active & lat).view(-1)).view(-1).to(torch.int32)
        plain = torch.zeros_like(d)
        _ccc_list(terma, density, electron, spacing, origin, iso, mu, ga, ca, ta, self.model,
                  lat_idx, plain)
        dmax = torch.clamp(d.max(), min=1e-12)
        bad = torch.zeros_like(active)
        bad.view(-1)[lat_idx.long()] = (d.view(-1)[lat_idx.long()] - plain.view(-1)[lat_idx.long()]).abs() > tol * dmax
        # Lattice cells (stride^3 voxels) whose checked voxel failed, dilated by
        # one cell, and cells with active voxels but no active checked voxel
        # (field edges), go to the plain CCCS.
        pool = lambda v: F.max_pool3d(v.float()[None, None], kernel_size=stride, stride=stride,
                                      ceil_mode=True)
        cells = F.max_pool3d(pool(bad), kernel_size=3, stride=1, padding=1)
        cells = torch.maximum(cells, pool(active) * (1.0 - pool(active & lat)))
        fb = F.interpolate(cells, scale_factor=stride, mode="nearest")[0, 0, :d.shape[0], :d.shape[1], :d.shape[2]] > 0
        fb = fb & active
        if ev:
            ev[2].record()
        fb_idx = torch.nonzero(fb.view(-1)).view(-1).to(torch.int32)
        _ccc_list(terma, density, electron, spacing, origin, iso, mu, ga, ca, ta, self.model, fb_idx, d)
        if ev:
            ev[3].record()
            torch.cuda.synchronize()
            timings.update(net_ms=ev[0].elapsed_time(ev[1]), check_ms=ev[1].elapsed_time(ev[2]),
                           fallback_ms=ev[2].elapsed_time(ev[3]), total_ms=ev[0].elapsed_time(ev[3]),
                           fallback_voxels=int(fb_idx.numel()), lattice_voxels=int(lat_idx.numel()),
                           active_voxels=int(active.sum().item()))
        d[~active] = 0.0
        return d


def _ccc_list(terma, density, electron, spacing, origin, iso, mu, ga, ca, ta, model, idx, out):
    keep: list = []
    m = engine.photon_args(model, keep)
    _lib.check(_lib.load().dh_stage_ccc_list_dev(
        terma.device.index or 0, _lib.grid(terma.shape, spacing, origin),
        _lib.photon_cp(iso, mu, ga, ca, ta), m, _lib.ptr(terma), _lib.ptr(density),
        _lib.ptr(electron), _lib.ptr(idx), int(idx.numel()), _lib.ptr(out), _st()))


def gamma(eval_dose: torch.Tensor, ref_dose: torch.Tensor, spacing, dta_mm=2.0, dd=0.02,
          cutoff=0.1, radius_vox=None):
    """Global gamma (dd of max reference dose, dta in mm) on the voxel grid;
    returns (pass rate over voxels with ref >= cutoff * max, gamma volume)."""
    ref = ref_dose.contiguous().float()
    ev = eval_dose.contiguous().float()
    dmax = float(ref.max().item())
    out = torch.empty_like(ref)
    rad = radius_vox if radius_vox is not None else int(math.ceil(2.0 * dta_mm / spacing))
    nz, ny, nx = ref.shape
    _lib.check(_lib.load().dh_gamma_dev(_lib.ptr(ev), _lib.ptr(ref), nz, ny, nx, float(spacing),
                                        float(dta_mm), float(dd * dmax), float(cutoff * dmax), rad,
                                        _lib.ptr(out), _st()))
    m = out >= 0
    n = int(m.sum().item())
    rate = float(((out <= 1.0) & m).sum().item()) / max(n, 1)
    return rate, out
