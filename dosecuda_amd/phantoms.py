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
