"""Vectorised NumPy restatement of the reference kernels for tiny grids.

TEST INFRASTRUCTURE ONLY (never imported by dosecuda_amd/).  It is a second,
independent transcription of the same formulae as oracle/dose_oracle.c,
written array-at-a-time instead of voxel-at-a-time, used to cross-check the C
oracle on small non-cubic grids where an axis-order or sign slip shows up.
FP32 decision arithmetic is emulated with numpy float32 operations
(fmaf as a float64 product-sum rounded once to float32).

  raytrace  CudaClasses.cu:180-228      smooth  IMPTClasses.cu:139-202
  pencil    IMPTClasses.cu:204-271      terma   IMRTClasses.cu:209-250
  ccc       IMRTClasses.cu:252-350
"""
from __future__ import annotations

import math

import numpy as np

F = np.float32
PI_F = F(3.141592654)


def fmaf(a, b, c):
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(F)


class G:
    def __init__(self, shape, s, origin, iso, gantry, couch, collimator, D):
        self.nz, self.ny, self.nx = shape
        self.s = F(s)
        self.inv_s = F(1.0) / self.s
        self.iso = (np.asarray(iso, F) - np.asarray(origin, F)).astype(F)
        adj = F(math.fmod(float(F(gantry + 180.0)), 360.0))
        ga = adj * PI_F / F(180.0)
        ta = F(couch) * PI_F / F(180.0)
        ca = F(collimator) * PI_F / F(180.0)
        self.sg, self.cg = F(np.sin(np.float64(ga))), F(np.cos(np.float64(ga)))
        self.st, self.ct = F(np.sin(np.float64(ta))), F(np.cos(np.float64(ta)))
        self.sc, self.cc = F(np.sin(np.float64(ca))), F(np.cos(np.float64(ca)))
        D = F(D)
        xg = -D * self.sg
        yg = D * self.cg
        self.src = np.array([xg * self.ct, yg, -xg * self.st], F)

    def set_trig(self, sg, cg, st, ct, sc, cc):
        """Adopt the C oracle's libm sin/cos so decisions compare bit-for-bit."""
        self.sg, self.cg, self.st, self.ct, self.sc, self.cc = (F(v) for v in (sg, cg, st, ct, sc, cc))

    def voxels(self):
        k, j, i = np.meshgrid(np.arange(self.nz), np.arange(self.ny), np.arange(self.nx), indexing="ij")
        x = i.astype(F) * self.s - self.iso[0]
        y = j.astype(F) * self.s - self.iso[1]
        z = k.astype(F) * self.s - self.iso[2]
        return x, y, z

    def to_head(self, x, y, z):
        sx, cx = -self.st, self.ct
        xt = x * cx + z * sx; yt = y; zt = -x * sx + z * cx
        sx, cx = -self.sg, self.cg
        xg = xt * cx - yt * sx; yg = xt * sx + yt * cx; zg = zt
        sx, cx = -self.sc, self.cc
        xc = xg * cx + zg * sx; yc = yg; zc = -xg * sx + zg * cx
        return -xc, zc, yc

    def to_img(self, hx, hy, hz):
        xz, yz, zz = -hx, hz, hy
        sx, cx = self.sc, self.cc
        xc = xz * cx + zz * sx; yc = yz; zc = -xz * sx + zz * cx
        sx, cx = self.sg, self.cg
        xg = xc * cx - yc * sx; yg = xc * sx + yc * cx; zg = zc
        sx, cx = self.st, self.ct
        return xg * cx + zg * sx, yg, -xg * sx + zg * cx


def sample(vol, tx, ty, tz, border):
    """Trilinear with border (float64 weights from float32 coordinates)."""
    nz, ny, nx = vol.shape
    u = np.asarray(tx, np.float64) - 0.5
    v = np.asarray(ty, np.float64) - 0.5
    w = np.asarray(tz, np.float64) - 0.5
    i0, j0, k0 = np.floor(u), np.floor(v), np.floor(w)
    fx, fy, fz = u - i0, v - j0, w - k0
    i0, j0, k0 = i0.astype(np.int64), j0.astype(np.int64), k0.astype(np.int64)
    out = np.zeros(np.shape(u))
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                ii, jj, kk = i0 + di, j0 + dj, k0 + dk
                ok = (ii >= 0) & (ii < nx) & (jj >= 0) & (jj < ny) & (kk >= 0) & (kk < nz)
                val = np.where(ok, vol[np.clip(kk, 0, nz - 1), np.clip(jj, 0, ny - 1),
                                       np.clip(ii, 0, nx - 1)].astype(np.float64), border)
                wt = (fx if di else 1 - fx) * (fy if dj else 1 - fy) * (fz if dk else 1 - fz)
                out += wt * val
    return out


def uvec(g, x, y, z):
    dx, dy, dz = g.src[0] - x, g.src[1] - y, g.src[2] - z
    ss = dx * dx + dy * dy + dz * dz
    r = (1.0 / np.sqrt(ss.astype(np.float64))).astype(F)
    return dx * r, dy * r, dz * r


def raytrace(g, rho, border=0.0):
    x, y, z = g.voxels()
    ux, uy, uz = uvec(g, x, y, z)
    acc = np.zeros(x.shape)
    alive = np.ones(x.shape, bool)
    L = F(0.0)
    while alive.any():
        qx, qy, qz = fmaf(ux, L, x), fmaf(uy, L, y), fmaf(uz, L, z)
        tx = fmaf(g.inv_s, qx + g.iso[0], F(0.5))
        ty = fmaf(g.inv_s, qy + g.iso[1], F(0.5))
        tz = fmaf(g.inv_s, qz + g.iso[2], F(0.5))
        inside = (tx >= 0) & (tx < F(g.nx)) & (ty >= 0) & (ty < F(g.ny)) & (tz >= 0) & (tz < F(g.nz))
        alive &= inside
        d = sample(rho, tx, ty, tz, border)
        acc += np.where(alive, np.maximum(d, 0.0), 0.0)
        L = L + F(1.0)
    return -0.05 + 0.1 * acc


def smooth(g, wet32):
    x, y, z = g.voxels()
    hx, hy, hz = g.to_head(x, y, z)
    c = wet32.astype(F)
    s = c.astype(np.float64)
    n = np.ones(c.shape)
    for a in range(6):
        ang = F(a) * PI_F / F(3.0)
        sn, cs = F(np.sin(np.float64(ang))), F(np.cos(np.float64(ang)))
        for r in range(1, 10):
            dr = F(r)
            use = (dr < c * F(10.0))
            qx, qy, qz = g.to_img(hx + dr * cs, hy + dr * sn, hz)
            # roundf: half away from zero
            ii = np.sign((qx + g.iso[0]) / g.s) * np.floor(np.abs((qx + g.iso[0]) / g.s) + 0.5)
            jj = np.sign((qy + g.iso[1]) / g.s) * np.floor(np.abs((qy + g.iso[1]) / g.s) + 0.5)
            kk = np.sign((qz + g.iso[2]) / g.s) * np.floor(np.abs((qz + g.iso[2]) / g.s) + 0.5)
            ok = use & (ii >= 0) & (ii < g.nx) & (jj >= 0) & (jj < g.ny) & (kk >= 0) & (kk < g.nz)
            vals = c[np.clip(kk, 0, g.nz - 1).astype(int), np.clip(jj, 0, g.ny - 1).astype(int),
                     np.clip(ii, 0, g.nx - 1).astype(int)]
            s += np.where(ok, vals, 0.0)
            n += ok
    return s / n


def pencil(g, vsadx, vsady, wet32, spots, dvp, lut_d, lut_s, lut_i):
    """spots: rows (x, y, mu, id) already grouped into layers by the caller's
    (id-sorted) order; layer selection as importLayers."""
    x, y, z = g.voxels()
    hx, hy, hz = g.to_head(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))
    dist = np.sqrt((g.src[0] - x.astype(np.float64)) ** 2 + (g.src[1] - y.astype(np.float64)) ** 2
                   + (g.src[2] - z.astype(np.float64)) ** 2)
    wf = wet32.astype(F) * F(10.0)
    w = wf.astype(np.float64)
    ids = spots[:, 3].astype(np.int64)
    dose = np.zeros(wet32.shape)
    for e in np.unique(ids):
        if e < 0 or e >= lut_d.shape[0]:
            continue
        sel = spots[ids == e]
        r80 = F(dvp[e, 1])
        act = ~(wf > F(1.1) * r80)
        d = lut_d[e, :400].astype(np.float64)
        idd = np.interp(w, d, lut_i[e, :400].astype(np.float64))
        sms = np.interp(w, d, lut_s[e, :400].astype(np.float64))
        r80 = float(r80)
        A, B, Cc = (float(v) for v in dvp[e, 2:5])
        da = dist - w + 0.7 * r80
        st = A * da * da + B * da + Cc + sms
        wc = np.minimum(np.maximum(w, 0.1), r80 - 0.1)
        sh = (2.85 + 0.0014 * r80 * np.log(wc + 3.0) + 0.06 * wc - 7.4e-5 * wc ** 2
              - 0.22 * r80 / (wc - r80 - 5.0) ** 2)
        hw = (0.052 * np.log(1.13 + wc / (11.2 - 0.023 * r80))
              + 0.35 * (0.0017 * r80 ** 2 - r80) / ((r80 + 3.0) ** 2 - wc ** 2)
              - 1.61e-9 * wc * (r80 + 3.0) ** 2)
        sh = np.maximum(sh, 0.0)
        hw = np.clip(hw, 0.0, 0.9)
        sht = np.hypot(st, sh)
        P = (1 - hw) * idd / (2 * np.pi * st ** 2)
        H = hw * idd / (2 * np.pi * sht ** 2)
        acc = np.zeros(w.shape)
        for sx, sy, mu, _ in sel.astype(np.float64):
            t = np.array([sx / vsadx, sy / vsady, -1.0])
            rx, ry, rz = hx - sx, hy - sy, hz
            d2 = rx * rx + ry * ry + rz * rz - (t[0] * rx + t[1] * ry + t[2] * rz) ** 2 / (t @ t)
            acc += mu * (P * np.exp(-0.5 * d2 / st ** 2) + H * np.exp(-0.5 * d2 / sht ** 2))
        dose += np.where(act, acc, 0.0)
    return dose


def _erf(x):
    from scipy.special import erf
    return erf(x)


def terma(g, m, mlc, wet32):
    x, y, z = g.voxels()
    hx, hy, hz = g.to_head(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))
    sad, ssd, dm = m["sad"], m["ssd_s"], m["d_mlc"]
    dps, dss = sad - hz, ssd - hz
    rc = np.hypot(hx, hy) * dps / sad
    oaf = np.interp(rc, m["prof_r"], m["prof_i"])
    soft = np.interp(rc, m["prof_r"], m["prof_s"])

    def ht(Z, sig):
        ms = (sad - dm) / sad
        dv = (Z - hz) / (dm - hz)
        inv = 1.0 / (sig * np.sqrt(2.0))
        tr = np.zeros(hz.shape)
        for x1, x2, yo, yw in mlc.astype(np.float64):
            t1 = (x1 * ms - hx) * dv + hx
            t2 = (x2 * ms - hx) * dv + hx
            e1 = ((yo - 0.5 * yw) * ms - hy) * dv + hy
            e2 = ((yo + 0.5 * yw) * ms - hy) * dv + hy
            tr += 0.25 * (_erf(t2 * inv) - _erf(t1 * inv)) * (_erf(e2 * inv) - _erf(e1 * inv))
        return tr

    Tp, Ts = ht(sad, m["src_p"]), ht(ssd, m["src_s"])
    W = wet32.astype(np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        tr = np.fmin(1.0, Ts / Tp)
    t = np.zeros(W.shape)
    for mu, pw, sw in zip(m["mu_e"], m["pw"], m["sw"]):
        att = np.exp(-mu * W * soft)
        t += (1 - tr) * pw * att * (sad / dps) ** 2 + tr * sw * att * (ssd / dss)
    lam, dmax = m["el_att"], m["el_dmax"]
    el = np.fmax(0.0, (np.exp(-lam * W) - np.exp(-lam * dmax)) / (1 - np.exp(-lam * dmax)))
    T = oaf * ((1 - m["ssw"]) * Tp * t + m["ssw"] * Ts * t)
    E = m["esw"] * (0.4 + 0.3 * tr) * el * Tp
    return T, E


def ccc(g, m, mu, terma32, rho32, el32):
    x, y, z = g.voxels()
    p = np.stack([x, y, z]).astype(np.float64)
    tc = [fmaf(g.inv_s, a + g.iso[q], F(0.5)) for q, a in enumerate((x, y, z))]
    T0 = sample(terma32, *tc, 0.0)
    src = g.src.astype(np.float64)
    u = src[:, None, None, None] - p
    u /= np.sqrt((u * u).sum(0))
    a = src / np.linalg.norm(src)
    c = np.tensordot(a, u, 1)
    k = np.cross(a[:, None, None, None], u, axis=0)
    sp = F(g.s / F(10.0))
    K = m["kernel"]
    dose = np.zeros(x.shape)
    for r in range(6):
        th = F(K[r, 0]) * PI_F / F(180.0)
        A, am, B, bm = (float(v) for v in K[r, 1:5])
        sth, cth = math.sin(float(th)), math.cos(float(th))
        for azi in range(12):
            phi = float(F(azi) * PI_F / F(6.0))
            dh = np.array([sth * math.cos(phi), sth * math.sin(phi), cth])
            # head -> image in float64
            xz, yz, zz = -dh[0], dh[2], dh[1]
            xc = xz * g.cc + zz * g.sc; yc = yz; zc = -xz * g.sc + zz * g.cc
            xg = xc * g.cg - yc * g.sg; yg = xc * g.sg + yc * g.cg; zg = zc
            d = np.array([xg * g.ct + zg * g.st, yg, -xg * g.st + zg * g.ct], np.float64)
            kd = np.tensordot(d, k, 1)
            dd = np.cross(k, d[:, None, None, None], axis=0) + d[:, None, None, None] * c + k * (kd / (1 + c))
            Rp = np.zeros(x.shape); Rs = np.zeros(x.shape)
            L = F(K[r, 5])
            while L >= F(0.0):
                Lm = float(L * F(10.0))
                q = p + dd * Lm
                t = [(q[i] + g.iso[i]) / g.s + 0.5 for i in range(3)]
                T = sample(terma32, *t, 0.0)
                D = np.maximum(sample(rho32, *t, float(F(0.0012))), float(F(0.0012))) * float(sp)
                e = np.exp(-am * D)
                Rp = Rp * e + T * sth * (A / am ** 2) * (1 - e)
                Rs = Rs * (1 - bm * D) + T * D * sth * (B / bm)
                L = F(L - sp)
            dose += am * Rp + bm * Rs
    dose += el32.astype(np.float64)
    ok = ~(T0 <= 0.01) & ~np.isnan(dose) & (dose >= 0)
    return np.where(ok, m["mu_cal"] * dose * mu, 0.0)
