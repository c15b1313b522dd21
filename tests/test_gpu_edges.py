"""Edge cases of the gather kernels and the drop-in (GPU).

1. The 64-bit offset paths.  The ray trace and the cone superposition take
   32-bit offsets whenever the padded volume is under 2 GiB and a padded
   plane has fewer than 2^23 elements; a one-slice 2900 x 2900 grid (planes of
   2903^2 > 2^23 elements) takes the 64-bit paths.  Its centre holds a
   crop, everything around it the border value, so each stage on the big grid
   must give the crop's own result (32-bit paths) inside the crop: the
   samples beyond the crop read the border either way.  The positions are
   computed from different origins (last-ulp differences), hence the
   tolerances.
2. Degenerate grids (one voxel thick, a single voxel) against the FP64 oracle.
3. Empty work: a field without spots and a TERMA volume below the cone
   threshold everywhere (an empty cone list) give zero dose.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import rel_err
from tests.test_gpu_parity import TOL_DOSE, TOL_WET, _cuda, _hetero, _report, _torch

pytestmark = pytest.mark.gpu
F = np.float32


@pytest.fixture(autouse=True)
def _need_gpu(gpu_ok):
    return gpu_ok


def _embed(crop, big_shape, j0, i0, fill):
    big = np.full(big_shape, fill, F)
    big[:, j0:j0 + crop.shape[1], i0:i0 + crop.shape[2]] = crop
    return big


@pytest.mark.parametrize("ga", [0.0, 63.0, 200.0])
def test_raytrace_64bit_offsets_match_crop(ga):
    t = _torch()
    from dosecuda_amd import engine
    s = 2.5
    crop = _hetero((1, 60, 80), 21)
    crop[:, :2, :] = 0.0
    crop[:, -2:, :] = 0.0
    crop[:, :, :2] = 0.0
    crop[:, :, -2:] = 0.0
    big_shape, j0, i0 = (1, 2900, 2900), 1420, 1410
    o_crop = np.array([-100.0, -75.0, 0.0], F)
    o_big = (o_crop - np.array([i0 * s, j0 * s, 0.0])).astype(F)
    iso = np.array([-5.0, 3.0, 0.0], F)
    print(f"[case] raytrace 64-bit offsets ga={ga}", flush=True)
    a = engine.stage_raytrace(_cuda(t, crop), s, o_crop, iso, ga, 0.0, 1000.0, 0.0).cpu().numpy()
    big = _embed(crop, big_shape, j0, i0, 0.0)
    b = engine.stage_raytrace(_cuda(t, big), s, o_big, iso, ga, 0.0, 1000.0, 0.0).cpu().numpy()
    b = b[:, j0:j0 + 60, i0:i0 + 80]
    mx, _ = rel_err(b, a)
    print(f"[parity] raytrace 64-bit vs 32-bit: max {mx:.3e}", flush=True)
    assert float(np.abs(a).max()) > 1.0 and mx <= 1e-5


def test_ccc_64bit_offsets_match_crop(photon_model):
    t = _torch()
    from dosecuda_amd import engine
    from dosecuda_amd.beam_models import square_field_mlc
    pm = photon_model
    s, ga = 2.5, 30.0
    rho = _hetero((1, 48, 56), 22, 0.3, 1.6)
    o_crop = np.array([-70.0, -60.0, 0.0], F)
    iso = np.array([0.0, 0.0, 0.0], F)
    mlc = square_field_mlc(pm, 4.0, 4.0)
    wet = engine.stage_raytrace(_cuda(t, rho), s, o_crop, iso, ga, 0.0, F(pm.primary_source_distance), 0.0012)
    terma, el = engine.stage_terma(wet, s, o_crop, iso, 100.0, ga, 0.0, 0.0, pm, mlc)
    print("[case] ccc 64-bit offsets", flush=True)
    a = engine.stage_ccc(terma, _cuda(t, rho), el, s, o_crop, iso, 100.0, ga, 0.0, 0.0, pm).cpu().numpy()
    big_shape, j0, i0 = (1, 2900, 2900), 1430, 1425
    o_big = (o_crop - np.array([i0 * s, j0 * s, 0.0])).astype(F)
    bt = _embed(terma.cpu().numpy(), big_shape, j0, i0, 0.0)
    bd = _embed(rho, big_shape, j0, i0, F(0.0012))
    be = _embed(el.cpu().numpy(), big_shape, j0, i0, 0.0)
    b = engine.stage_ccc(_cuda(t, bt), _cuda(t, bd), _cuda(t, be), s, o_big, iso, 100.0, ga, 0.0, 0.0,
                         pm).cpu().numpy()
    assert np.count_nonzero(b[:, :j0]) == 0 and np.count_nonzero(b[:, j0 + 48:]) == 0
    b = b[:, j0:j0 + 48, i0:i0 + 56]
    mx, _ = rel_err(b, a)
    print(f"[parity] ccc 64-bit vs 32-bit: max {mx:.3e}", flush=True)
    assert float(a.max()) > 0.0 and mx <= 1e-5


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 5, 7), (6, 1, 4), (2, 2, 2)])
def test_degenerate_grids_raytrace_and_smooth(shape):
    t = _torch()
    from dosecuda_amd import engine
    s, ga, ta = 3.0, 40.0, 15.0
    rho = _hetero(shape, 23)
    origin = np.array([-2.0, 1.0, 0.5], F)
    iso = np.array([1.0, 2.0, 0.5], F)
    print(f"[case] degenerate raytrace/smooth shape={shape}", flush=True)
    got = engine.stage_raytrace(_cuda(t, rho), s, origin, iso, ga, ta, 2000.0, 0.0).cpu().numpy()
    g = O.geom(shape, s, origin, iso, ga, ta, 0.0, 2000.0)
    ref, _ = O.raytrace(g, rho, 0.0)
    _report("raytrace", got, ref, TOL_WET)
    sm = engine.stage_smooth(_cuda(t, got), s, origin, iso, ga, ta, 2000.0, 2000.0).cpu().numpy()
    _, ref_sm, _ = O.proton_wet(rho, s, origin, iso, ga, ta, 2000.0, 2000.0)
    assert np.all(np.isfinite(sm)) and rel_err(sm, ref_sm)[0] <= 1e-5


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 9, 11), (7, 1, 5)])
def test_degenerate_grids_photon_dropin(photon_model, shape):
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import VolumeObject, square_field_mlc
    pm = photon_model
    s = 2.5
    rho = _hetero(shape, 24, 0.5, 1.5)
    iso = np.array([(shape[2] - 1) * s / 2, (shape[1] - 1) * s / 2, (shape[0] - 1) * s / 2], F)
    mlc = square_field_mlc(pm, 3.0, 3.0)
    vol = VolumeObject(rho, (0.0, 0.0, 0.0), (s, s, s))

    class CP:
        pass
    cp = CP()
    cp.iso, cp.mu, cp.mlc, cp.ga, cp.ca, cp.ta = iso, 100.0, mlc, 90.0, 0.0, 0.0
    print(f"[case] degenerate photon drop-in shape={shape}", flush=True)
    got = dose_kernels.photon_dose_cuda(pm, vol, cp, 0)
    ref = O.photon_dose(rho, s, np.zeros(3), iso, 100.0, 90.0, 0.0, 0.0, pm, mlc)
    assert got.shape == shape and np.all(np.isfinite(got))
    if float(np.abs(ref).max()) > 0.0:
        _report("photon drop-in", got, ref, TOL_DOSE)
    else:
        assert np.all(got == 0.0)


def test_field_without_spots_gives_zero_dose(proton_model):
    _torch()
    from dosecuda_amd import dose_kernels
    from dosecuda_amd.beam_models import ProtonField, VolumeObject
    m = proton_model
    rho = _hetero((10, 12, 14), 25)
    vol = VolumeObject(rho, (0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    f = ProtonField(np.array([13.0, 11.0, 9.0], F), 0.0, 0.0, np.empty((0, 4), F))
    print("[case] proton field without spots", flush=True)
    wet = dose_kernels.proton_raytrace_cuda(m, vol, f, 0)
    w = VolumeObject(np.array(wet, F), (0.0, 0.0, 0.0), (2.0, 2.0, 2.0))
    dose = dose_kernels.proton_spot_cuda(m, vol, w, f, 0)
    assert dose.shape == rho.shape and np.all(dose == 0.0)


def test_empty_cone_list_gives_zero_dose(photon_model):
    t = _torch()
    from dosecuda_amd import engine
    rho = _hetero((12, 14, 16), 26)
    terma = np.full(rho.shape, 0.005, F)                 # below the 0.01 cone threshold everywhere
    el = np.ones(rho.shape, F)
    print("[case] empty cone list", flush=True)
    d = engine.stage_ccc(_cuda(t, terma), _cuda(t, rho), _cuda(t, el), 2.5, np.zeros(3, F),
                         np.array([10.0, 10.0, 10.0], F), 100.0, 0.0, 0.0, 0.0, photon_model).cpu().numpy()
    assert np.all(d == 0.0)
