"""Learned CCCS surrogate with plain-CCCS fallback.

A small 3-D convolutional encoder-decoder (U-Net, widths 16/32/64/128, about
1.5 M parameters) predicts the kernel-superposition part of the collapsed-
cone dose of one photon beam from its TERMA and density volumes; the
electron-contamination term and the MU calibration are added exactly, as the
plain CCCS adds them (IMRTClasses.cu:338-346).  The network works on a
half-resolution grid without losing input detail: each 2x2x2 block of
(TERMA, density) becomes 16 channels of one cell (space-to-depth), three more
channels carry the beam direction, and 8 output channels are the 2x2x2 dose
sub-voxels of the cell (depth-to-space).  A 128^3 volume is one 64^3 patch.
Every convolution (forward, input gradient and weight gradient) is a
hand-written BF16 MFMA kernel of libdosehip.so with FP32 accumulation; torch
supplies memory, pooling/upsampling/concatenation glue and the Adam update
of the FP32 master weights.

Trust, on the timed path: the plain CCCS (`k_ccc`) is evaluated at one
voxel of each lattice cell of `stride`^3 voxels (its low corner when that has
TERMA > 0.01, else its centre voxel); every cell whose surrogate value there
differs from the plain value by more than `tol` of the maximum, together with
its neighbour cells, and every cell with neither voxel active is recomputed
with the plain CCCS voxel by voxel.  The full plain
CCCS and the 2%/2 mm gamma index score it, off the timed path.

No reference counterpart: DoseCUDA computes only the plain CCCS
(IMRTClasses.cu:252-350).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np
import torch
import torch.nn.functional as F

from . import _lib, engine

SLOPE = 0.1          # LeakyReLU negative slope
IN_CH = 32           # 8 TERMA + 8 density sub-voxels, 3 beam direction, spacing / 2.5 mm, 12 zero
SP_REF = 2.5         # spacing channel = spacing / SP_REF
N_SUB = 8            # output sub-voxels per cell
AIR = 0.0012         # density of the padding outside the volume (IMRTClasses.cuh:6)


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


def conv3d(xa: torch.Tensor, wprep: torch.Tensor, bias, cout: int, ksize: int, act: bool,
           want_bf16=True, want_f32=False, xb=None, up=False, nco=0, ksplit=0):
    """Convolution of the channel concatenation [upsample(xa) if up else xa,
    xb] (xb optional) on channels-last bf16 tensors."""
    ca = xa.shape[-1]
    if xb is not None:
        n, d, h, w, cb = xb.shape
    else:
        n, d, h, w = xa.shape[:4]
        if up:
            d, h, w = 2 * d, 2 * h, 2 * w
        cb = 0
    y = torch.empty((n, d, h, w, cout), dtype=torch.bfloat16, device=xa.device) if want_bf16 else None
    y32 = torch.empty((n, d, h, w, cout), dtype=torch.float32, device=xa.device) if want_f32 else None
    _lib.check(_lib.load().dh_nn_conv3d_dev(_lib.ptr(xa), ca, 1 if up else 0, _lib.ptr(xb), cb, n, d, h, w,
                                            _lib.ptr(wprep), _lib.ptr(bias), cout, wprep.shape[1], ksize,
                                            1 if act else 0, SLOPE, _lib.ptr(y), _lib.ptr(y32), int(nco),
                                            int(ksplit), _st()))
    return y, y32


def pool2(x: torch.Tensor) -> torch.Tensor:
    n, d, h, w, c = x.shape
    y = torch.empty((n, d // 2, h // 2, w // 2, c), dtype=torch.bfloat16, device=x.device)
    _lib.check(_lib.load().dh_nn_pool2_dev(_lib.ptr(x), n, d, h, w, c, _lib.ptr(y), _st()))
    return y


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


class _Pool(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return pool2(x.contiguous())

    @staticmethod
    def backward(ctx, gy):
        g = gy.float() * 0.125
        g = g.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)
        return g.to(torch.bfloat16)


def _up(x):
    return x.repeat_interleave(2, 1).repeat_interleave(2, 2).repeat_interleave(2, 3)


class SurrogateNet(torch.nn.Module):
    """U-Net on channels-last bf16 tensors [N, D, H, W, IN_CH] (D, H, W
    multiples of 8); output fp32 [N, D, H, W, 16], of which channels 0-7 are
    the 2x2x2 sub-voxel values (normalised superposition dose)."""

    def __init__(self, widths=(16, 32, 64, 128), seed=1004):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        a, b, c, d = widths
        self.widths = tuple(widths)
        self.e0 = torch.nn.ModuleList([Conv(IN_CH, a, gen=g), Conv(a, a, gen=g)])
        self.e1 = torch.nn.ModuleList([Conv(a, b, gen=g), Conv(b, b, gen=g)])
        self.e2 = torch.nn.ModuleList([Conv(b, c, gen=g), Conv(c, c, gen=g)])
        self.bt = torch.nn.ModuleList([Conv(c, d, gen=g), Conv(d, d, gen=g)])
        self.d2 = torch.nn.ModuleList([Conv(d + c, c, gen=g), Conv(c, c, gen=g)])
        self.d1 = torch.nn.ModuleList([Conv(c + b, b, gen=g), Conv(b, b, gen=g)])
        self.d0 = torch.nn.ModuleList([Conv(b + a, a, gen=g), Conv(a, a, gen=g)])
        # 1x1 head with 16 outputs (a multiple of 16 keeps the input-gradient
        # convolution on whole MFMA K slices); channels 0-7 are used
        self.head = Conv(a, 16, ksize=1, act=False, out32=True, gen=g)
        # target normalisation: superposition dose / (max TERMA * out_scale)
        self.register_buffer("out_scale", torch.ones(()))

    @staticmethod
    def _seq(mods, x):
        for m in mods:
            x = m(x)
        return x

    def forward(self, x):
        """Training forward (autograd; decoder concatenations materialised)."""
        self._wf = None
        s0 = self._seq(self.e0, x)
        s1 = self._seq(self.e1, _Pool.apply(s0))
        s2 = self._seq(self.e2, _Pool.apply(s1))
        t = self._seq(self.bt, _Pool.apply(s2))
        t = self._seq(self.d2, torch.cat([_up(t), s2], dim=-1))
        t = self._seq(self.d1, torch.cat([_up(t), s1], dim=-1))
        t = self._seq(self.d0, torch.cat([_up(t), s0], dim=-1))
        return self.head(t)

    def prepare(self):
        """Cache the bf16 forward-layout weights for infer()."""
        self._wf = {id(m): prep_weights(m.w.detach(), 0) for m in self.modules() if isinstance(m, Conv)}
        return self

    @torch.no_grad()
    def infer(self, x):
        """Inference forward with the cached weights: the decoder's
        upsample + concatenation is read in place by the convolution kernel.
        Same kernels and K split as forward(), so the same result."""
        if getattr(self, "_wf", None) is None:
            self.prepare()

        def c(m, xa, xb=None, up=False):
            y, y32 = conv3d(xa, self._wf[id(m)], m.b, m.w.shape[2], m.ksize, m.act, want_bf16=not m.out32,
                            want_f32=m.out32, xb=xb, up=up)
            return y32 if m.out32 else y

        s0 = c(self.e0[1], c(self.e0[0], x))
        s1 = c(self.e1[1], c(self.e1[0], pool2(s0)))
        s2 = c(self.e2[1], c(self.e2[0], pool2(s1)))
        t = c(self.bt[1], c(self.bt[0], pool2(s2)))
        t = c(self.d2[1], c(self.d2[0], t, s2, up=True))
        t = c(self.d1[1], c(self.d1[0], t, s1, up=True))
        t = c(self.d0[1], c(self.d0[0], t, s0, up=True))
        return c(self.head, t)

    def n_params(self):
        return sum(p.numel() for n, p in self.named_parameters())

    def flops(self, cells):
        """FLOPs (2 per multiply-add) of one forward pass over `cells` cells
        of the half-resolution grid: 2 * taps * cin * cout per output cell."""
        tot = 0.0
        for mods, scale in ((self.e0, 1), (self.e1, 8), (self.e2, 64), (self.bt, 512),
                            (self.d2, 64), (self.d1, 8), (self.d0, 1)):
            for m in mods:
                t, ci, co = m.w.shape
                tot += 2.0 * t * ci * co * cells / scale
        t, ci, co = self.head.w.shape
        return tot + 2.0 * t * ci * co * cells


def save(net: SurrogateNet, path: str, compact: bool = False):
    """safetensors; compact = convolution weights as bf16 (what inference
    uses: the bf16 rounding of the stored values is exact, so a compact
    file gives the same inference results) and the rest fp32."""
    from safetensors.torch import save_file
    sd = {}
    for k, v in net.state_dict().items():
        v = v.detach().cpu().contiguous()
        sd[k] = v.to(torch.bfloat16) if (compact and k.endswith(".w")) else v.float()
    save_file(sd, path, metadata={"widths": ",".join(str(w) for w in net.widths),
                                  "in_ch": str(IN_CH), "n_sub": str(N_SUB)})


def load(path: str, device="cuda") -> SurrogateNet:
    from safetensors import safe_open
    from safetensors.torch import load_file
    with safe_open(path, framework="pt") as f:
        meta = f.metadata() or {}
    if int(meta.get("in_ch", IN_CH)) != IN_CH or int(meta.get("n_sub", N_SUB)) != N_SUB:
        raise ValueError(f"{path}: surrogate weights for another input/output layout ({meta})")
    widths = tuple(int(w) for w in meta.get("widths", "16,32,64,128").split(","))
    net = SurrogateNet(widths)
    net.load_state_dict({k: v.float() for k, v in load_file(path).items()})
    return net.to(device)


# ---------------------------------------------------------------------------
# inputs / targets
# ---------------------------------------------------------------------------
def beam_direction(gantry, couch):
    """Unit vector source -> isocentre in volume (x, y, z) axes, with the
    engine's gantry and couch conventions (CudaClasses.cu:27-49)."""
    ga = math.radians(math.fmod(gantry + 180.0, 360.0))
    ta = math.radians(couch)
    xg, yg = -math.sin(ga), math.cos(ga)
    src = np.array([xg * math.cos(ta), yg, -xg * math.sin(ta)])
    return (-src / np.linalg.norm(src)).astype(np.float32)


def padded_shape(shape):
    return tuple(_ceil(int(s), 16) for s in shape)


def _pad(v, shape, value):
    D, H, W = v.shape
    if (D, H, W) == tuple(shape):
        return v
    return F.pad(v, (0, shape[2] - W, 0, shape[1] - H, 0, shape[0] - D), value=value)


def _s2d(v):
    """(D, H, W) -> (D/2, H/2, W/2, 8), sub-voxel (dz, dy, dx) -> 4 dz + 2 dy + dx."""
    D, H, W = v.shape
    return v.view(D // 2, 2, H // 2, 2, W // 2, 2).permute(0, 2, 4, 1, 3, 5).reshape(D // 2, H // 2, W // 2, 8)


def _d2s(v):
    d, h, w, _ = v.shape
    return v.reshape(d, h, w, 2, 2, 2).permute(0, 3, 1, 4, 2, 5).reshape(2 * d, 2 * h, 2 * w)


def make_input(terma: torch.Tensor, density: torch.Tensor, direction, spacing, out=None):
    """fp32 (D, H, W) TERMA and density -> channels-last bf16 input
    [D'/2, H'/2, W'/2, IN_CH] (D', H', W' = sizes padded to multiples of 16)
    and the TERMA scale (its maximum)."""
    shp = padded_shape(terma.shape)
    scale = torch.clamp(terma.amax(), min=1e-12)
    x = out if out is not None else torch.empty(tuple(s // 2 for s in shp) + (IN_CH,), dtype=torch.bfloat16,
                                                device=terma.device)
    x[..., 0:8] = _s2d(_pad(terma / scale, shp, 0.0))
    x[..., 8:16] = _s2d(_pad(density, shp, AIR))
    x[..., 16:19] = torch.as_tensor(direction, dtype=torch.float32, device=terma.device)
    x[..., 19] = float(np.float32(spacing) / np.float32(SP_REF))
    x[..., 20:] = 0
    return x, scale


def dose_scale(model, mu):
    return float(np.float32(model.mu_calibration)) * float(mu)


def superposition_target(dose, electron, terma, model, mu):
    """Invert the plain CCCS's output formula: the kernel-superposition sum
    per voxel (zero where TERMA <= 0.01 or where the plain dose was clamped)."""
    sup = dose / dose_scale(model, mu) - electron
    return torch.where((terma > 0.01) & (dose > 0), sup, torch.zeros_like(sup))


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


def _lung_bone_phantom(n, spacing, rng: np.random.Generator, device="cuda"):
    """Density of a phantom of the 128^3 lung/bone plan's family (the
    geometry of phantoms.imrt_lung_bone_case: water, a -700 HU lung ellipsoid
    with semi-axes 30-50 mm within 20 mm of the centre, a +1000 HU bone
    cylinder along z of radius 10-20 mm within 40 mm; bone over lung), built
    on the device."""
    from .beam_models import density_from_hu
    d_lung, d_water, d_bone = (float(v) for v in density_from_hu(np.array([-700.0, 0.0, 1000.0])))
    c = (n - 1) / 2.0 * spacing
    z = torch.arange(n, device=device, dtype=torch.float32)[:, None, None] * spacing
    y = torch.arange(n, device=device, dtype=torch.float32)[None, :, None] * spacing
    x = torch.arange(n, device=device, dtype=torch.float32)[None, None, :] * spacing
    lc = c + rng.uniform(-20.0, 20.0, 3)
    ax = rng.uniform(30.0, 50.0, 3)
    br = float(rng.uniform(10.0, 20.0))
    bc = c + rng.uniform(-40.0, 40.0, 2)
    rho = torch.full((n, n, n), d_water, device=device)
    lung = ((x - lc[0]) / ax[0]) ** 2 + ((y - lc[1]) / ax[1]) ** 2 + ((z - lc[2]) / ax[2]) ** 2 <= 1.0
    rho[lung.expand(n, n, n)] = d_lung
    bone = (x - bc[0]) ** 2 + (y - bc[1]) ** 2 <= br * br
    rho[bone.expand(n, n, n)] = d_bone
    return rho.contiguous()


def generate(model, n_samples, n=128, spacing=(2.0, 3.0), seed=1004, log=None, family="random"):
    """Training samples from the plain CCCS: random phantoms, voxel spacings
    (uniform in the `spacing` range, or fixed), gantry angles, isocentre
    shifts of up to 2 cm and square fields of 3-20 cm.  Returns inputs
    X [S, n/2, n/2, n/2, 16] bf16 (TERMA/max and density sub-voxels), the
    per-sample channels [S, 4] (beam direction, spacing / 2.5 mm) and targets
    Y [S, n/2, n/2, n/2, 8] fp32 (superposition dose / max TERMA).
    family "mix": every other sample is a lung/bone phantom of the plan
    family at 2.5 mm (drawn from this generator's own seeded stream, not the
    evaluated phantom's) with a 10 x 10 cm field half of the time."""
    from .beam_models import square_field_mlc
    rng = np.random.default_rng(seed)
    h = n // 2
    X = torch.empty((n_samples, h, h, h, 16), dtype=torch.bfloat16, device="cuda")
    Y = torch.empty((n_samples, h, h, h, N_SUB), dtype=torch.float32, device="cuda")
    dirs = torch.empty((n_samples, 4), dtype=torch.float32, device="cuda")
    terma = torch.empty((n, n, n), device="cuda")
    el = torch.empty_like(terma)
    dose = torch.empty_like(terma)
    t0 = time.time()
    for s in range(n_samples):
        sp = float(rng.uniform(*spacing)) if isinstance(spacing, (tuple, list)) else float(spacing)
        fs = float(rng.uniform(3.0, 20.0))
        if family == "mix" and s % 2 == 1:
            sp = 2.5
            rho = _lung_bone_phantom(n, sp, rng)
            if rng.random() < 0.5:
                fs = 10.0
        else:
            rho = random_phantom(n, sp, rng)
        ga = float(rng.uniform(0.0, 360.0))
        iso = (np.full(3, (n - 1) / 2 * sp) + rng.uniform(-20.0, 20.0, 3)).astype(np.float32)
        mlc = square_field_mlc(model, fs, fs)
        engine.photon_dose(rho, sp, np.zeros(3), iso, 100.0, ga, 0.0, 0.0, model, mlc,
                           out=dose, terma_out=terma, electron_out=el)
        scale = torch.clamp(terma.amax(), min=1e-12)
        X[s, ..., 0:8] = _s2d(terma / scale)
        X[s, ..., 8:16] = _s2d(rho)
        dirs[s, :3] = torch.from_numpy(beam_direction(ga, 0.0))
        dirs[s, 3] = float(np.float32(sp) / np.float32(SP_REF))
        Y[s] = _s2d(superposition_target(dose, el, terma, model, 100.0) / scale)
        if log and (s + 1) % max(1, n_samples // 4) == 0:
            log(f"generated {s + 1}/{n_samples} samples ({time.time() - t0:.1f} s)")
    return X, dirs, Y


def _edge_weight(Y, alpha):
    """Per-target-voxel loss weight 1 + alpha * |grad Y| / max Y (sub-voxel
    layout in and out): high-gradient voxels (build-up, penumbra) count more."""
    B = Y.shape[0]
    out = torch.empty_like(Y)
    for b in range(B):
        v = _d2s(Y[b])
        g = torch.zeros_like(v)
        for ax in range(3):
            d = torch.diff(v, dim=ax)
            pad = [0, 0] * 3
            pad[2 * (2 - ax) + 1] = 1
            g = g + torch.nn.functional.pad(d, pad).abs()
        m = torch.clamp(v.amax(), min=1e-12)
        out[b] = _s2d(1.0 + alpha * g / m * 10.0)
    return out


def assemble(X, chans):
    """[B, d, h, w, 16] + [B, 4] (direction, spacing channel) -> network
    input [B, d, h, w, IN_CH]."""
    B = X.shape[0]
    x = torch.zeros(X.shape[:4] + (IN_CH,), dtype=torch.bfloat16, device=X.device)
    x[..., 0:16] = X
    x[..., 16:20] = chans.view(B, 1, 1, 1, 4).to(torch.bfloat16)
    return x


@torch.no_grad()
def eval_loss(net: SurrogateNet, X, chans, Y, batch=4):
    """The training loss (weighted mean squared error, in target units) of
    the network on samples X / chans / Y from generate()."""
    tot, n = 0.0, 0
    for i in range(0, X.shape[0], batch):
        xb = assemble(X[i:i + batch], chans[i:i + batch])
        yb = Y[i:i + batch]
        pred = net(xb)[..., :N_SUB] * net.out_scale
        wgt = torch.where(yb > 0, 1.0, 0.05)
        tot += float((wgt * (pred - yb) ** 2).sum())
        n += yb.numel()
    return tot / max(n, 1)


def train(net: SurrogateNet, model, n_samples=20000, chunk=1000, passes=3, batch=4, lr=5e-4,
          n=128, spacing=(2.0, 3.0), seed=1004, log=None, fit_scale=True, edge=0.0, family="random"):
    """Online training: the plain CCCS generates `chunk` samples at a time,
    each chunk is used for `passes` epochs, `n_samples` in total.  Adam on the
    FP32 master weights, one-cycle learning rate; loss = mean squared error
    over voxels of the field (target > 0) plus a small weight outside.
    fit_scale=False keeps the network's output scale (continued training);
    edge > 0 weights the in-field voxels by 1 + edge * 10 |grad target| / max
    (per voxel of the 2.5 mm grid); family as in generate()."""
    n_chunks = max(1, n_samples // chunk)
    steps_per_chunk = max(1, chunk * passes // batch)
    total = n_chunks * steps_per_chunk
    # beta2 0.99: the second moment follows changes of the gradient scale
    # (phantoms of different spacing) faster than the default 0.999
    opt = torch.optim.Adam(net.parameters(), lr=lr, betas=(0.9, 0.99))
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=total, pct_start=0.05)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    t0 = time.time()
    step = 0
    tgen = 0.0
    run, skipped = None, 0
    best, good, lr_scale, restores = None, None, 1.0, 0
    for ci in range(n_chunks):
        tg = time.time()
        X, dirs, Y = generate(model, chunk, n=n, spacing=spacing, seed=seed * 1000 + ci, family=family)
        torch.cuda.synchronize()
        tgen += time.time() - tg
        peak = Y.amax(dim=(1, 2, 3, 4))
        if ci == 0 and fit_scale:
            net.out_scale.fill_(float(peak[peak > 0].mean()) if bool((peak > 0).any()) else 1.0)
        if log:
            log(f"chunk {ci}: target peak mean {float(peak.mean()):.3g} max {float(peak.max()):.3g}, "
                f"min TERMA-normalised input max {float(X[..., 0:8].float().amax(dim=(1, 2, 3, 4)).min()):.3g}")
        inv = 1.0 / float(net.out_scale)
        for _ in range(steps_per_chunk):
            idx = torch.randint(0, chunk, (batch,), generator=g).cuda()
            xb = assemble(X[idx], dirs[idx])
            yb = Y[idx] * inv
            pred = net(xb)[..., :N_SUB]
            wgt = torch.where(yb > 0, _edge_weight(yb, edge) if edge > 0 else 1.0, 0.05)
            loss = (wgt * (pred - yb) ** 2).mean()
            opt.zero_grad(set_to_none=True)
            loss.backward()
            lv = float(loss.item())
            # A non-finite loss or a spike above 20x the running level skips
            # the update (gradients are clipped to unit norm otherwise).  The
            # level follows every finite loss, capped at 20x itself, so it
            # adapts when the loss level itself moves.
            if math.isfinite(lv) and (run is None or lv < 20.0 * run):
                torch.nn.utils.clip_grad_norm_(net.parameters(), 1.0)
                opt.step()
            else:
                skipped += 1
            if math.isfinite(lv):
                run = lv if run is None else 0.98 * run + 0.02 * min(lv, 20.0 * run)
            # divergence guard: keep a snapshot while the running level is
            # near its best; when it climbs 10x above the best, restore the
            # snapshot, reset Adam's moments and halve the learning rate
            if run is not None and step % 250 == 249:
                if best is None or run < 2.0 * best:
                    best = run if best is None else min(best, run)
                    good = {k: v.detach().clone() for k, v in net.state_dict().items()}
                elif run > 10.0 * best and good is not None:
                    net.load_state_dict(good)
                    opt.state.clear()
                    lr_scale *= 0.5
                    restores += 1
                    run = best
                    if log:
                        log(f"step {step + 1}: divergence, weights restored, learning rate x{lr_scale}")
            sched.step()
            if lr_scale != 1.0:
                for gp in opt.param_groups:
                    gp["lr"] *= lr_scale
            step += 1
            if step % max(1, total // 40) == 0 or step == total:
                hist.append(run if run is not None else lv)
                if log:
                    log(f"step {step}/{total} loss {hist[-1]:.3e} lr {opt.param_groups[0]['lr']:.2e} "
                        f"skipped {skipped} restored {restores} ({time.time() - t0:.1f} s, generation {tgen:.1f} s)")
        del X, dirs, Y
    return hist


# ---------------------------------------------------------------------------
# inference with the plain fallback
# ---------------------------------------------------------------------------
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "cccs_surrogate.safetensors")
# lattice stride, tolerance (of Dmax) and dilation of the check the benchmark
# and the plan-level tests use with the shipped weights: on the 7-beam 128^3
# lung/bone plan every beam of the whole path passes 2%/2 mm at >= 99.5%
DEFAULT_CHECK = (4, 0.015, 1)


class NativeNet:
    """Handle of the library's native forward pass over a network's cached
    bf16 weights (dh_sur_net_create); keeps those buffers alive."""

    def __init__(self, net: SurrogateNet, device=None):
        net.prepare()
        convs = [*net.e0, *net.e1, *net.e2, *net.bt, *net.d2, *net.d1, *net.d0, net.head]
        self.weights = [net._wf[id(m)] for m in convs]
        self.biases = [m.b.detach().contiguous() for m in convs]
        dev = self.weights[0].device.index or 0 if device is None else int(device)
        W = (C.c_void_p * len(convs))(*[w.data_ptr() for w in self.weights])
        B = (C.c_void_p * len(convs))(*[b.data_ptr() for b in self.biases])
        widths = (C.c_int32 * 4)(*net.widths)
        h = C.c_void_p()
        _lib.check(_lib.load().dh_sur_net_create(dev, widths, IN_CH, W, B, SLOPE, float(net.out_scale), C.byref(h)))
        self.handle = h
        self.device = dev

    def close(self):
        if self.handle:
            _lib.load().dh_sur_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SurrogateCCCS:
    """Surrogate photon dose of one beam from the plain path's TERMA,
    electron-contamination and density volumes, checked against the plain
    CCCS on a lattice of voxels with the plain CCCS as per-voxel fallback.
    Everything runs on the device without host synchronisation, so a beam
    (or a plan of beams) can be captured into a HIP graph."""

    def __init__(self, net: SurrogateNet, model, stride=3, tol=0.01, dilate=1):
        self.net = net.eval().prepare()
        self.out_scale = float(net.out_scale)
        self.model = model
        self.stride, self.tol, self.dilate = int(stride), float(tol), int(dilate)
        self._buf = {}
        self.native = NativeNet(self.net)

    def _buffers(self, shape, device, stride):
        dev = torch.device(device)
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
        key = (tuple(int(v) for v in shape), idx, stride)
        b = self._buf.get(key)
        if b is None:
            pz, py, px = padded_shape(shape)
            nw = C.c_int64()
            _lib.check(_lib.load().dh_sur_check_work_size(_lib.grid(shape, 1.0, np.zeros(3)), stride, C.byref(nw)))
            b = dict(x=torch.empty((1, pz // 2, py // 2, px // 2, IN_CH), dtype=torch.bfloat16, device=device),
                     tmax=torch.zeros(1, device=device), dmax=torch.zeros(1, device=device),
                     work=torch.zeros(int(nw.value), dtype=torch.int32, device=device),
                     plain=torch.empty(tuple(shape), dtype=torch.float32, device=device))
            self._buf[key] = b
        return b

    @torch.no_grad()
    def predict(self, terma, density, electron, gantry, couch, mu, spacing, out=None, stride=None):
        """Surrogate dose, finished like the plain CCCS's output: MU
        calibration, electron term, zero outside TERMA > 0.01, no negatives."""
        nz, ny, nx = terma.shape
        b = self._buffers(terma.shape, terma.device, self.stride if stride is None else int(stride))
        out = torch.empty_like(terma) if out is None else out
        d = beam_direction(gantry, couch)
        lib = _lib.load()
        _lib.check(lib.dh_sur_pack_dev(nz, ny, nx, _lib.ptr(terma), _lib.ptr(density), _lib.ptr(d), float(spacing),
                                       _lib.ptr(b["x"]), _lib.ptr(b["tmax"]), _st()))
        y = self.net.infer(b["x"])
        _lib.check(lib.dh_sur_finish_dev(nz, ny, nx, _lib.ptr(y), _lib.ptr(b["tmax"]), self.out_scale,
                                         _lib.ptr(terma), _lib.ptr(electron), dose_scale(self.model, mu),
                                         _lib.ptr(out), _lib.ptr(b["dmax"]), _st()))
        return out

    @torch.no_grad()
    def dose(self, terma, density, electron, spacing, origin, iso, mu, ga, ca, ta, mlc=None,
             stride=None, tol=None, out=None, timings=None):
        """Checked surrogate dose of one control point (see class doc) in one
        call of the library (dh_sur_dose_dev).  timings (a dict) receives the
        total time and the lattice and fallback voxel counts."""
        stride = self.stride if stride is None else int(stride)
        tol = self.tol if tol is None else float(tol)
        out = torch.empty_like(terma) if out is None else out
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)] if timings is not None else None
        if ev:
            ev[0].record()
        keep: list = []
        m = engine.photon_args(self.model, keep)
        _lib.check(_lib.load().dh_sur_dose_dev(
            self.native.handle, _lib.grid(terma.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta), m,
            _lib.ptr(terma), _lib.ptr(density), _lib.ptr(electron), stride, tol, self.dilate, _lib.ptr(out), _st()))
        if ev:
            ev[1].record()
            torch.cuda.synchronize()
            lat, fb = C.c_int32(), C.c_int32()
            _lib.check(_lib.load().dh_sur_counts(terma.device.index or 0, C.byref(lat), C.byref(fb)))
            timings.update(total_ms=ev[0].elapsed_time(ev[1]), lattice_voxels=int(lat.value),
                           fallback_voxels=int(fb.value), active_voxels=int((~(terma <= 0.01)).sum().item()))
        return out

    @torch.no_grad()
    def dose_batch(self, termas, density, electrons, spacing, origin, cps, stride=None, tol=None, outs=None):
        """Checked surrogate doses of several control points of one grid and
        density in one call of the library (dh_sur_dose_batch_dev): cps is a
        list of (iso, mu, ga, ca, ta), termas / electrons the matching
        volumes.  The points share network launches (up to 8 per batch);
        each dose equals dose() of that point alone, bit for bit."""
        stride = self.stride if stride is None else int(stride)
        tol = self.tol if tol is None else float(tol)
        n = len(cps)
        if n == 0 or len(termas) != n or len(electrons) != n:
            raise ValueError("need one TERMA and one electron volume per control point")
        outs = [torch.empty_like(t) for t in termas] if outs is None else list(outs)
        if len(outs) != n:
            raise ValueError("need one output volume per control point")
        for v in (*termas, *electrons, *outs):
            if v.shape != density.shape or v.dtype != torch.float32 or not v.is_contiguous():
                raise ValueError("volumes must be contiguous float32 of the density's shape")
        nbytes = density.numel() * 4
        ins = [(v.data_ptr(), v.data_ptr() + nbytes) for v in (*termas, *electrons, density)]
        outr = [(v.data_ptr(), v.data_ptr() + nbytes) for v in outs]
        for k, (a, b) in enumerate(outr):
            if any(a < d and c < b for c, d in ins) or any(a < d and c < b for c, d in outr[k + 1:]):
                raise ValueError("output volumes must not overlap each other or the input volumes")
        arr = (_lib.PhotonCP * n)(*[_lib.photon_cp(*cp) for cp in cps])
        P = lambda vs: (C.c_void_p * n)(*[v.data_ptr() for v in vs])
        keep: list = []
        m = engine.photon_args(self.model, keep)
        _lib.check(_lib.load().dh_sur_dose_batch_dev(
            self.native.handle, _lib.grid(density.shape, spacing, origin), n, arr, m, P(termas),
            _lib.ptr(density), P(electrons), stride, tol, self.dilate, P(outs), _st()))
        return outs

    @torch.no_grad()
    def plan_doses(self, density, spacing, origin, cps, mlcs, outs=None, work=None, step=0.0):
        """Checked surrogate doses of a plan's control points on one grid: the
        points' TERMA and electron volumes (engine.photon_terma_batch; step
        0: the depth from the beam's-eye-view grid, > 0: a ray trace sampled
        every step mm), then dose_batch.  cps: (iso, mu, ga, ca, ta) per
        point, mlcs: their apertures.  work: optional lists (termas,
        electrons) of volumes to reuse."""
        termas, els = work if work is not None else ([torch.empty_like(density) for _ in cps],
                                                     [torch.empty_like(density) for _ in cps])
        engine.photon_terma_batch(density, spacing, origin, cps, self.model, mlcs, step=step,
                                  terma_outs=termas, electron_outs=els)
        return self.dose_batch(termas, density, els, spacing, origin, cps, outs=outs)

    @torch.no_grad()
    def dose_torch(self, terma, density, electron, spacing, origin, iso, mu, ga, ca, ta, mlc=None,
                   stride=None, tol=None, out=None, timings=None):
        """The same checked dose with the network driven from Python
        (SurrogateNet.infer) and the check through dh_sur_check_dev."""
        stride = self.stride if stride is None else int(stride)
        tol = self.tol if tol is None else float(tol)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)] if timings is not None else None
        if ev:
            ev[0].record()
        d = self.predict(terma, density, electron, ga, ta, mu, spacing, out=out, stride=stride)
        if ev:
            ev[1].record()
        b = self._buffers(terma.shape, terma.device, stride)
        keep: list = []
        m = engine.photon_args(self.model, keep)
        _lib.check(_lib.load().dh_sur_check_dev(
            terma.device.index or 0, _lib.grid(terma.shape, spacing, origin), _lib.photon_cp(iso, mu, ga, ca, ta),
            m, _lib.ptr(terma), _lib.ptr(density), _lib.ptr(electron), _lib.ptr(d), _lib.ptr(b["dmax"]), stride,
            tol, self.dilate, _lib.ptr(b["work"]), b["work"].numel(), _lib.ptr(b["plain"]), _st()))
        if ev:
            ev[2].record()
            torch.cuda.synchronize()
            cnt = b["work"][:2].tolist()
            timings.update(net_ms=ev[0].elapsed_time(ev[1]), check_ms=ev[1].elapsed_time(ev[2]),
                           total_ms=ev[0].elapsed_time(ev[2]), lattice_voxels=int(cnt[0]),
                           fallback_voxels=int(cnt[1]), active_voxels=int((~(terma <= 0.01)).sum().item()))
        return d

    def counts(self, shape, device="cuda", stride=None):
        """(lattice voxels, fallback voxels) of the last checked dose."""
        b = self._buffers(shape, device, self.stride if stride is None else int(stride))
        c = b["work"][:2].tolist()
        return int(c[0]), int(c[1])


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


# ---------------------------------------------------------------------------
# evaluation on the 7-beam lung/bone phantom and the command line
# ---------------------------------------------------------------------------
def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def time_plan_path(net: SurrogateNet, model, n_beams=7, reps=10, check=None):
    """The whole checked surrogate path of the 128^3 lung/bone plan (7
    coplanar 10x10 cm beams): `plan_doses` from the density captured as one
    HIP graph and replayed `reps` times; returns the milliseconds per beam and
    each beam's gamma (2%/2 mm, 10% cut) against the plain CCCS.  Only this
    path runs between the capture and the gamma checks, so a kernel trace of
    a process that calls it shows the path's own launches."""
    from .beam_models import square_field_mlc
    from .phantoms import imrt_beams, imrt_lung_bone_case, to_density
    st, tol, dil = check or DEFAULT_CHECK
    case = imrt_lung_bone_case(seed=1002, n=128, spacing=2.5)
    rho = torch.from_numpy(to_density(case)).cuda()
    mlc = square_field_mlc(model, 10.0, 10.0)
    sur = SurrogateCCCS(net, model, stride=st, tol=tol, dilate=dil)
    angles = imrt_beams(n_beams)
    cps = [(case.iso, 100.0, ga, 0.0, 0.0) for ga in angles]
    mlcs = [mlc] * len(angles)
    work = ([torch.empty_like(rho) for _ in angles], [torch.empty_like(rho) for _ in angles])
    outs = [torch.empty_like(rho) for _ in angles]
    sur.plan_doses(rho, case.spacing, case.origin, cps, mlcs, outs=outs, work=work)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        sur.plan_doses(rho, case.spacing, case.origin, cps, mlcs, outs=outs, work=work)
    g.replay()
    torch.cuda.synchronize()
    ms = _time(lambda: g.replay(), reps) / len(angles)
    plains = [engine.photon_dose(rho, case.spacing, case.origin, case.iso, 100.0, ga, 0.0, 0.0, model, mlc)
              for ga in angles]
    gm = [gamma(o, p, case.spacing)[0] for o, p in zip(outs, plains)]
    return dict(ms_per_beam=ms, beam_gamma=gm, min_gamma=min(gm), beams=len(angles), reps=reps,
                check=dict(stride=st, tol=tol, dilate=dil))


def evaluate(net: SurrogateNet, model, n_beams=7, stride=3, tol=0.01, reps=5, log=None, dilate=1):
    """Per-beam and plan gamma (2%/2 mm, 10% cut) of the surrogate alone and
    of the checked surrogate against the plain CCCS on the 128^3 lung/bone
    phantom with 10x10 cm fields, and the timings of both paths (eager, and
    the checked surrogate of the whole plan replayed as one HIP graph)."""
    from .beam_models import square_field_mlc
    from .phantoms import imrt_beams, imrt_lung_bone_case, to_density
    case = imrt_lung_bone_case(seed=1002, n=128, spacing=2.5)
    rho = torch.from_numpy(to_density(case)).cuda()
    mlc = square_field_mlc(model, 10.0, 10.0)
    sur = SurrogateCCCS(net, model, stride=stride, tol=tol, dilate=dilate)
    sp, org, iso = case.spacing, case.origin, case.iso
    angles = imrt_beams(n_beams)
    res = dict(beams=[], stride=stride, tol=tol, dilate=dilate)
    sums = {k: torch.zeros_like(rho) for k in ("plain", "net", "checked")}
    termas, els, plains, checked = [], [], [], []
    for ga in angles:
        terma, el = torch.empty_like(rho), torch.empty_like(rho)
        plain = engine.photon_dose(rho, sp, org, iso, 100.0, ga, 0.0, 0.0, model, mlc,
                                   terma_out=terma, electron_out=el)
        t_plain = _time(lambda: engine.photon_dose(rho, sp, org, iso, 100.0, ga, 0.0, 0.0, model, mlc,
                                                   terma_out=terma, electron_out=el), reps)
        engine.prof_reset()
        engine.prof_enable(True)
        for _ in range(reps):
            engine.photon_dose(rho, sp, org, iso, 100.0, ga, 0.0, 0.0, model, mlc,
                               terma_out=terma, electron_out=el)
        torch.cuda.synchronize()
        engine.prof_enable(False)
        stage = {k: engine.prof_get(k)[0] / reps for k in ("raytrace", "terma", "pack_tr", "ccc_compact", "ccc",
                                                           "ccc_finish")}
        t_net = _time(lambda: sur.predict(terma, rho, el, ga, 0.0, 100.0, sp), reps)
        dn = sur.predict(terma, rho, el, ga, 0.0, 100.0, sp)
        tm = {}
        dc = sur.dose(terma, rho, el, sp, org, iso, 100.0, ga, 0.0, 0.0, timings=tm)
        t_chk = _time(lambda: sur.dose(terma, rho, el, sp, org, iso, 100.0, ga, 0.0, 0.0), reps)
        t_chk_torch = _time(lambda: sur.dose_torch(terma, rho, el, sp, org, iso, 100.0, ga, 0.0, 0.0), reps)
        engine.prof_reset()
        engine.prof_enable(True)
        for _ in range(reps):
            sur.dose(terma, rho, el, sp, org, iso, 100.0, ga, 0.0, 0.0)
        torch.cuda.synchronize()
        engine.prof_enable(False)
        sstage = {k: engine.prof_get(k)[0] / reps for k in ("sur_pack", "nn_conv3d", "sur_finish", "pack_tr",
                                                           "sur_lattice", "ccc_list", "ccc_finish",
                                                           "sur_fallback")}
        gn, _ = gamma(dn, plain, sp)
        gc, _ = gamma(dc, plain, sp)
        pm = float(plain.max())
        b = dict(gantry=ga, gamma_net=gn, gamma_checked=gc,
                 max_err_net=float((dn - plain).abs().max()) / pm,
                 max_err_checked=float((dc - plain).abs().max()) / pm,
                 rms_err_checked=float((dc - plain).pow(2).mean().sqrt()) / pm,
                 plain_beam_ms=t_plain, plain_stage_ms=stage, surrogate_net_ms=t_net,
                 surrogate_checked_ms=t_chk, surrogate_checked_torch_ms=t_chk_torch,
                 surrogate_stage_ms=sstage, checked_split=tm)
        res["beams"].append(b)
        if log:
            log(json.dumps(b))
        sums["plain"] += plain
        sums["net"] += dn
        sums["checked"] += dc
        termas.append(terma)
        els.append(el)
        plains.append(plain)
        checked.append(dc)
    # the checked surrogate of the whole plan as one HIP graph
    outs = [torch.empty_like(rho) for _ in angles]

    def plan():
        for i, ga in enumerate(angles):
            sur.dose(termas[i], rho, els[i], sp, org, iso, 100.0, ga, 0.0, 0.0, out=outs[i])

    plan()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan()
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        res["graph_replay_equal"] = all(bool(torch.equal(o, c)) for o, c in zip(outs, checked))
        if not res["graph_replay_equal"]:
            break
    res["graph_single_ms_per_beam"] = _time(lambda: g.replay(), reps) / len(angles)
    # the same plan with the beams batched through the network (one launch
    # per layer for all beams); must equal the per-beam doses bit for bit
    cps = [(iso, 100.0, ga, 0.0, 0.0) for ga in angles]
    outs_b = [torch.empty_like(rho) for _ in angles]

    def plan_batch():
        sur.dose_batch(termas, rho, els, sp, org, cps, outs=outs_b)

    plan_batch()
    torch.cuda.synchronize()
    res["batch_equal_single"] = all(bool(torch.equal(o, c)) for o, c in zip(outs_b, checked))
    eager_b = [o.clone() for o in outs_b]
    gb = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gb):
        plan_batch()
    for _ in range(3):
        gb.replay()
        torch.cuda.synchronize()
        res["graph_replay_equal"] = res["graph_replay_equal"] and all(
            bool(torch.equal(o, c)) for o, c in zip(outs_b, eager_b))
    res["graph_checked_ms_per_beam"] = _time(lambda: gb.replay(), reps) / len(angles)
    # the whole surrogate path of the plan from the density: TERMA with one
    # ray sample per voxel spacing, then the checked surrogate of all beams
    # batched, as one HIP graph; judged by gamma against the plain plan
    mlcs = [mlc for _ in angles]
    work = ([torch.empty_like(rho) for _ in angles], [torch.empty_like(rho) for _ in angles])
    outs_p = [torch.empty_like(rho) for _ in angles]
    sur.plan_doses(rho, sp, org, cps, mlcs, outs=outs_p, work=work)
    torch.cuda.synchronize()
    eager_p = [o.clone() for o in outs_p]
    gp = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gp):
        sur.plan_doses(rho, sp, org, cps, mlcs, outs=outs_p, work=work)
    for _ in range(3):
        gp.replay()
        torch.cuda.synchronize()
        res["graph_replay_equal"] = res["graph_replay_equal"] and all(
            bool(torch.equal(o, c)) for o, c in zip(outs_p, eager_p))
    res["path_ms_per_beam"] = _time(lambda: gp.replay(), reps) / len(angles)
    pg = [gamma(o, p, sp)[0] for o, p in zip(outs_p, plains)]
    res["path_min_gamma"] = min(pg)
    res["path_plan_gamma"] = gamma(sum(outs_p), sums["plain"], sp)[0]
    res["path_beam_gamma"] = pg
    # the same path with the depth from a ray trace sampled once per voxel spacing
    step = max(1.0, float(np.float32(sp)))
    gs = torch.cuda.CUDAGraph()
    sur.plan_doses(rho, sp, org, cps, mlcs, outs=outs_p, work=work, step=step)
    torch.cuda.synchronize()
    with torch.cuda.graph(gs):
        sur.plan_doses(rho, sp, org, cps, mlcs, outs=outs_p, work=work, step=step)
    gs.replay()
    torch.cuda.synchronize()
    res["path_raytrace_ms_per_beam"] = _time(lambda: gs.replay(), reps) / len(angles)
    res["path_raytrace_min_gamma"] = min(gamma(o, p, sp)[0] for o, p in zip(outs_p, plains))
    engine.prof_reset()
    engine.prof_enable(True)
    for _ in range(reps):
        sur.plan_doses(rho, sp, org, cps, mlcs, outs=outs_p, work=work)
    torch.cuda.synchronize()
    engine.prof_enable(False)
    res["path_stage_ms_per_beam"] = {k: engine.prof_get(k)[0] / reps / len(angles)
                                     for k in ("pad", "bev_prefix", "raytrace", "terma", "nn_conv3d", "ccc_list",
                                               "sur_pack", "sur_finish", "sur_lattice", "sur_fallback")}
    engine.prof_reset()
    engine.prof_enable(True)
    for _ in range(reps):
        plan_batch()
    torch.cuda.synchronize()
    engine.prof_enable(False)
    res["batch_conv_ms_per_beam"] = engine.prof_get("nn_conv3d")[0] / reps / len(angles)
    res["plan_gamma_net"] = gamma(sums["net"], sums["plain"], sp)[0]
    res["plan_gamma_checked"] = gamma(sums["checked"], sums["plain"], sp)[0]
    B = res["beams"]
    mean = lambda k: float(np.mean([b[k] for b in B]))
    sup_stage = float(np.mean([b["plain_stage_ms"]["ccc"] + b["plain_stage_ms"]["ccc_finish"] +
                               b["plain_stage_ms"]["ccc_compact"] + b["plain_stage_ms"]["pack_tr"] for b in B]))
    res.update(plain_beam_ms=mean("plain_beam_ms"), plain_superposition_ms=sup_stage,
               surrogate_net_ms=mean("surrogate_net_ms"), surrogate_checked_ms=mean("surrogate_checked_ms"),
               min_gamma_net=min(b["gamma_net"] for b in B), min_gamma_checked=min(b["gamma_checked"] for b in B),
               mean_fallback_fraction=float(np.mean([b["checked_split"]["fallback_voxels"] /
                                                     max(1, b["checked_split"]["active_voxels"]) for b in B])),
               speedup_superposition=sup_stage / res["graph_checked_ms_per_beam"],
               net_gflop=net.flops(64 ** 3) / 1e9, params=net.n_params())
    return res


def main(argv=None):
    import argparse
    from .beam_models import PhotonBeamModel
    ap = argparse.ArgumentParser(prog="python -m dosecuda_amd.surrogate")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="train on plain-CCCS samples and save the weights")
    t.add_argument("--samples", type=int, default=20000)
    t.add_argument("--chunk", type=int, default=1000)
    t.add_argument("--passes", type=int, default=3)
    t.add_argument("--batch", type=int, default=4)
    t.add_argument("--lr", type=float, default=5e-4)
    t.add_argument("--seed", type=int, default=1004)
    t.add_argument("--out", default=WEIGHTS)
    t.add_argument("--init", default=None, help="continue from these weights (keeps their output scale)")
    t.add_argument("--eval", action="store_true", help="evaluate after training")
    t.add_argument("--edge", type=float, default=0.0, help="gradient weighting of the loss (0 off)")
    t.add_argument("--family", default="random", choices=["random", "mix"],
                   help="phantoms: random, or every other one of the lung/bone plan family")
    e = sub.add_parser("eval", help="gamma and timing on the 7-beam lung/bone phantom")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--sweep", default=None, help="stride:tol[:dilate] settings, e.g. 4:0.01,4:0.02:1")
    for q in (t, e):
        q.add_argument("--json", default=None, help="write the evaluation as JSON here")
        q.add_argument("--stride", type=int, default=3)
        q.add_argument("--tol", type=float, default=0.01)
        q.add_argument("--dilate", type=int, default=1, help="0 none, 1 face neighbours, 2 all 26")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    model = PhotonBeamModel()
    if a.cmd == "train":
        net = load(a.init) if a.init else SurrogateNet().cuda()
        log(f"surrogate: {net.n_params()} parameters, {net.flops(64 ** 3) / 1e9:.1f} GFLOP per 64^3 patch")
        train(net, model, n_samples=a.samples, chunk=a.chunk, passes=a.passes, batch=a.batch, lr=a.lr,
              seed=a.seed, log=log, fit_scale=not a.init, edge=a.edge, family=a.family)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        save(net, a.out, compact=True)
        log(f"saved {a.out}")
        if not a.eval:
            return
    else:
        net = load(a.weights)
    # convolution variants of the 64^3-cell forward, as the benchmark picks
    # them (timed over the 7-beam plan's batch)
    from . import autotune
    autotune.tune_surrogate(net, (64, 64, 64), reps=5, batch=7)
    if a.cmd == "eval" and a.sweep:
        out = []
        for pair in a.sweep.split(","):
            f = pair.split(":")
            r = evaluate(net, model, stride=int(f[0]), tol=float(f[1]), dilate=int(f[2]) if len(f) > 2 else 1)
            summary = {k: v for k, v in r.items() if k != "beams"}
            log(json.dumps(summary))
            out.append(r)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(out, f, indent=1)
        return
    r = evaluate(net, model, stride=a.stride, tol=a.tol, dilate=a.dilate, log=log)
    summary = {k: v for k, v in r.items() if k != "beams"}
    log(json.dumps(summary))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(r, f, indent=1)


if __name__ == "__main__":
    main()
