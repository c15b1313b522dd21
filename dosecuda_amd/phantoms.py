"""Seeded synthetic phantoms and fields for the benchmark workloads.

Shapes are voxelised with a voxel-centre-in-shape test in float64 from a few
seeded parameters, so the phantom does not depend on numpy's RNG stream
beyond those parameters.  Voxel centres sit at i * spacing (origin 0) and the
isocentre at the volume centre.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .beam_models import (PhotonBeamModel, ProtonBeamModel, density_from_hu, rlsp_from_hu,
                          square_field_mlc)


@dataclass
class Case:
    hu: np.ndarray
    spacing: float
    origin: np.ndarray
    iso: np.ndarray
    meta: dict = field(default_factory=dict)


def _centres(shape, s):
    nz, ny, nx = shape
    z = np.arange(nz) * s
    y = np.arange(ny) * s
    x = np.arange(nx) * s
    return z[:, None, None], y[None, :, None], x[None, None, :]


def _iso(shape, s):
    nz, ny, nx = shape
    return np.array([(nx - 1) / 2.0 * s, (ny - 1) / 2.0 * s, (nz - 1) / 2.0 * s], dtype=np.float32)


def impt_slab_case(seed=1001, n=100, spacing=2.5, slab_mm=20.0):
    """Water cube with a +1000 HU slab across x-z at a seeded depth 30-70 mm
    from the upstream (j = 0) face; gantry 0 beams enter there."""
    rng = np.random.default_rng(seed)
    depth = float(rng.uniform(30.0, 70.0))
    shape = (n, n, n)
    hu = np.zeros(shape, np.float32)
    z, y, x = _centres(shape, spacing)
    sl = (y >= depth) & (y < depth + slab_mm)
    hu[np.broadcast_to(sl, shape)] = 1000.0
    return Case(hu, spacing, np.zeros(3, np.float32), _iso(shape, spacing),
                dict(slab_depth_mm=depth, seed=seed))


def impt_spot_grid(model: ProtonBeamModel, seed=1001, n_side=20, pitch=5.0, n_layers=10,
                   e_lo=100.0, e_hi=160.0):
    """20x20 spots at 5 mm pitch over n_layers energies drawn without
    replacement from the model's energies in [e_lo, e_hi] MeV, unit MU."""
    rng = np.random.default_rng(seed + 7)
    ids = np.nonzero((model.energy_labels >= e_lo) & (model.energy_labels <= e_hi))[0]
    chosen = np.sort(rng.choice(ids, size=n_layers, replace=False))
    c = (np.arange(n_side) - (n_side - 1) / 2.0) * pitch
    xx, yy = np.meshgrid(c, c, indexing="xy")
    rows = []
    for e in chosen:
        rows.append(np.column_stack([xx.ravel(), yy.ravel(), np.ones(xx.size), np.full(xx.size, e)]))
    return np.array(np.vstack(rows), dtype=np.single), chosen


def imrt_lung_bone_case(seed=1002, n=128, spacing=2.5):
    """Water cube with a lung ellipsoid (-700 HU, semi-axes 30-50 mm) and a
    bone cylinder along z (+1000 HU, radius 10-20 mm); bone overrides lung."""
    rng = np.random.default_rng(seed)
    shape = (n, n, n)
    iso = _iso(shape, spacing).astype(np.float64)
    lc = iso + rng.uniform(-20.0, 20.0, 3)
    ax = rng.uniform(30.0, 50.0, 3)
    br = float(rng.uniform(10.0, 20.0))
    bc = iso[:2] + rng.uniform(-40.0, 40.0, 2)
    z, y, x = _centres(shape, spacing)
    hu = np.zeros(shape, np.float32)
    lung = ((x - lc[0]) / ax[0]) ** 2 + ((y - lc[1]) / ax[1]) ** 2 + ((z - lc[2]) / ax[2]) ** 2 <= 1.0
    hu[np.broadcast_to(lung, shape)] = -700.0
    bone = (x - bc[0]) ** 2 + (y - bc[1]) ** 2 <= br * br
    hu[np.broadcast_to(bone, shape)] = 1000.0
    return Case(hu, spacing, np.zeros(3, np.float32), iso.astype(np.float32),
                dict(seed=seed, lung_centre=lc.tolist(), lung_axes=ax.tolist(),
                     bone_radius=br, bone_centre=bc.tolist()))


def imrt_beams(n_beams=7):
    """Coplanar gantry angles 360*b/n (0, 51.43, ..., 308.57 for 7)."""
    return [360.0 * b / n_beams for b in range(n_beams)]


def spheres_case(seed=1003, shape=(160, 256, 256), spacing=2.0, n_min=6, n_max=10):
    """Water with 6-10 spheres, radius 10-40 mm, HU uniform in [-800, 1200]."""
    rng = np.random.default_rng(seed)
    hu = np.zeros(shape, np.float32)
    z, y, x = _centres(shape, spacing)
    ext = np.array([shape[2], shape[1], shape[0]]) * spacing
    spheres = []
    for _ in range(int(rng.integers(n_min, n_max + 1))):
        r = float(rng.uniform(10.0, 40.0))
        c = rng.uniform(r, ext - r)
        h = float(rng.uniform(-800.0, 1200.0))
        m = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
        hu[np.broadcast_to(m, shape)] = h
        spheres.append((c.tolist(), r, h))
    return Case(hu, spacing, np.zeros(3, np.float32), _iso(shape, spacing),
                dict(seed=seed, spheres=spheres))


def to_rlsp(case: Case):
    return rlsp_from_hu(case.hu)


def to_density(case: Case):
    return density_from_hu(case.hu)


def default_photon_field(model: PhotonBeamModel, size_cm=10.0):
    return square_field_mlc(model, size_cm, size_cm)


def influence_fields(model: ProtonBeamModel, seed=1003, gantries=(0.0, 120.0, 240.0),
                     layers_per_field=(8, 8, 9), id_max=85, n_min=200, n_max=900, total=15000,
                     half_width=60.0):
    """Spot lists of the batched influence workload: 25 layers split 8/8/9
    over gantry 0/120/240, energy ids from 0..id_max (70.2-197.6 MeV), spots
    per layer U{200..900} rescaled to about `total`, positions uniform in
    +-half_width mm (BEV), MU 1.  Returns [(gantry, spot_rows)]."""
    rng = np.random.default_rng(seed)
    n_layers = int(sum(layers_per_field))
    ids = np.sort(rng.choice(np.arange(id_max + 1), size=n_layers, replace=False))
    counts = rng.integers(n_min, n_max + 1, n_layers).astype(np.float64)
    counts = np.maximum(1, np.round(counts * total / counts.sum())).astype(int)
    order = rng.permutation(n_layers)
    out, pos = [], 0
    for g, nl in zip(gantries, layers_per_field):
        rows = []
        for li in order[pos:pos + nl]:
            n = int(counts[li])
            xy = rng.uniform(-half_width, half_width, (n, 2))
            rows.append(np.column_stack([xy, np.ones(n), np.full(n, ids[li])]))
        pos += nl
        out.append((float(g), np.array(np.vstack(rows), dtype=np.single)))
    return out


@dataclass
class MixedCase:
    """One case of the mixed workload: a phantom (water with spheres) and
    either proton fields (spot rows per gantry) or photon beams (control
    points: gantry, MU, MLC rows)."""
    kind: str
    shape: tuple
    spacing: float
    spheres: list
    fields: list = field(default_factory=list)      # proton: [(gantry, spot_rows)]
    cps: list = field(default_factory=list)         # photon: [(gantry, mu, mlc_rows)]


def mixed_cases(pmodel: ProtonBeamModel, xmodel: PhotonBeamModel, seed=1005, n_cases=64,
                sides=(96, 256), spacing=(2.0, 3.0), spots=(2000, 8000), fields=(1, 3), beams=(5, 9),
                segments=(5, 20)):
    """The mixed batched workload: n_cases phantoms with per-axis sides
    uniform in `sides` voxels and spacing uniform in `spacing` mm, water with
    2-5 spheres (radius 10-40 mm, HU uniform in [-800, 1200]); even cases
    are IMPT (1-3 fields at random gantry angles, 2k-8k unit-MU spots each
    over 6-15 energy layers whose R80 fits in 70% of the smallest extent,
    positions uniform in +-45 mm BEV), odd cases IMRT (5-9 equally spaced
    beams from a random start angle, 5-20 MLC segments each: the leaf pairs
    within a random y window open on a random x interval, MU 5-40)."""
    rng = np.random.default_rng(seed)
    r80 = np.asarray(pmodel.divergence_params)[:, 1]
    offs = np.asarray(xmodel.mlc_offsets, np.float64)
    widths = np.asarray(xmodel.mlc_widths, np.float64)
    out = []
    for c in range(n_cases):
        shape = tuple(int(v) for v in rng.integers(sides[0], sides[1] + 1, 3))
        s = float(rng.uniform(*spacing))
        ext = np.array([shape[2], shape[1], shape[0]], np.float64) * s
        sph = []
        for _ in range(int(rng.integers(2, 6))):
            r = float(rng.uniform(10.0, min(40.0, 0.25 * ext.min())))
            sph.append((rng.uniform(r, ext - r).tolist(), r, float(rng.uniform(-800.0, 1200.0))))
        mc = MixedCase("impt" if c % 2 == 0 else "imrt", shape, s, sph)
        if mc.kind == "impt":
            ok = np.nonzero((r80 >= 30.0) & (r80 <= 0.7 * ext.min()))[0]
            if ok.size == 0:
                ok = np.array([int(np.argmin(r80))])
            for _ in range(int(rng.integers(fields[0], fields[1] + 1))):
                g = float(rng.uniform(0.0, 360.0))
                n = int(rng.integers(spots[0], spots[1] + 1))
                nl = int(min(ok.size, rng.integers(6, 16)))
                ids = np.sort(rng.choice(ok, size=nl, replace=False))
                per = rng.multinomial(n - nl, np.full(nl, 1.0 / nl)) + 1
                rows = [np.column_stack([rng.uniform(-45.0, 45.0, (k, 2)), np.ones(k), np.full(k, e)])
                        for e, k in zip(ids, per)]
                mc.fields.append((g, np.array(np.vstack(rows), dtype=np.single)))
        else:
            nb = int(rng.integers(beams[0], beams[1] + 1))
            g0 = float(rng.uniform(0.0, 360.0))
            for b in range(nb):
                g = float((g0 + 360.0 * b / nb) % 360.0)
                for _ in range(int(rng.integers(segments[0], segments[1] + 1))):
                    ylo = float(rng.uniform(-60.0, 20.0))
                    yhi = ylo + float(rng.uniform(10.0, 60.0))
                    mlc = np.zeros((offs.size, 4), np.float32)
                    mlc[:, 2] = offs
                    mlc[:, 3] = widths
                    for i in np.nonzero((offs >= ylo) & (offs <= yhi))[0]:
                        x1 = float(rng.uniform(-60.0, 20.0))
                        mlc[i, 0] = x1
                        mlc[i, 1] = min(x1 + float(rng.uniform(5.0, 50.0)), 70.0)
                    mc.cps.append((g, float(rng.uniform(5.0, 40.0)), mlc))
        out.append(mc)
    return out


def voxelize_mixed(case: MixedCase, device="cpu"):
    """HU volume of a mixed case (torch float32 on `device`)."""
    import torch
    nz, ny, nx = case.shape
    s = case.spacing
    z = torch.arange(nz, dtype=torch.float64, device=device)[:, None, None] * s
    y = torch.arange(ny, dtype=torch.float64, device=device)[None, :, None] * s
    x = torch.arange(nx, dtype=torch.float64, device=device)[None, None, :] * s
    hu = torch.zeros(case.shape, dtype=torch.float32, device=device)
    for c, r, h in case.spheres:
        m = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r * r
        hu[m] = h
    return hu


def mixed_iso(case: MixedCase):
    return _iso(case.shape, case.spacing)
