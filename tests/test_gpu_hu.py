"""HU -> RLSP / density on the device against np.interp (GPU).

dh_hu_convert_dev replaces the plan layer's host conversions
(plan_impt.py:112-122, plan_imrt.py:114-124: np.interp over the machine's HU
table, then float32).  np.interp is the specification here, so the checks are
bit-for-bit against it: random HU, every table node exactly, values beyond
both ends, NaN, a table with repeated nodes and a one-node table, float32 and
float64 input, and the machines' own tables.
"""
import numpy as np
import pytest
import torch

from dosecuda_amd import beam_models as bm
from dosecuda_amd import engine

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _need_gpu(gpu_ok):
    return gpu_ok


def _ref(hu, xp, fp):
    return np.array(np.interp(hu, xp, fp), dtype=np.single)


def _hu_values(xp, n, seed, dtype):
    rng = np.random.default_rng(seed)
    v = np.concatenate([rng.uniform(xp[0] - 500, xp[-1] + 500, n), xp, np.nextafter(xp, -np.inf),
                        np.nextafter(xp, np.inf), [xp[0] - 1e6, xp[-1] + 1e6, np.nan, -np.inf, np.inf]])
    return v.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("kind,machine", [("protons", "HitachiProbeatJHU"), ("photons", "VarianTrueBeamHF"),
                                          ("photons", "ElektaVersaSibAxesse")])
def test_machine_tables_equal_np_interp(dtype, kind, machine):
    print(f"[case] hu table {kind}/{machine} {np.dtype(dtype).name}", flush=True)
    xp, fp = bm.hu_table(kind, machine)
    hu = _hu_values(xp, 200_000, 5, dtype)
    got = engine.hu_convert(torch.from_numpy(hu).cuda(), xp, fp).cpu().numpy()
    ref = _ref(hu, xp, fp)
    assert np.array_equal(got, ref, equal_nan=True)


def test_repeated_and_single_node_tables():
    print("[case] hu table repeated / single nodes", flush=True)
    xp = np.array([-1000.0, -200.0, -200.0, 0.0, 0.0, 0.0, 1500.0, 3000.0])
    fp = np.array([0.001, 0.8, 0.85, 1.0, 1.02, 1.03, 1.8, 2.5])
    hu = _hu_values(xp, 100_000, 6, np.float32)
    got = engine.hu_convert(torch.from_numpy(hu).cuda(), xp, fp).cpu().numpy()
    assert np.array_equal(got, _ref(hu, xp, fp), equal_nan=True)
    hu = np.array([-5.0, 0.0, 7.0], np.float32)
    got = engine.hu_convert(torch.from_numpy(hu).cuda(), [3.0], [0.25]).cpu().numpy()
    assert np.array_equal(got, np.full(3, 0.25, np.float32))


def test_volume_and_errors():
    print("[case] hu table volume + errors", flush=True)
    xp, fp = bm.hu_table("protons", "HitachiProbeatJHU")
    hu = np.random.default_rng(7).uniform(-1100, 2500, (37, 41, 43)).astype(np.float32)
    got = engine.hu_convert(torch.from_numpy(hu).cuda(), xp, fp).cpu().numpy()
    assert got.shape == hu.shape and np.array_equal(got, bm.rlsp_from_hu(hu))
    with pytest.raises(ValueError, match="1..128"):
        engine.hu_convert(torch.zeros(4, device="cuda"), np.arange(200.0), np.arange(200.0))
    with pytest.raises(ValueError):
        engine.hu_convert(torch.zeros(4, device="cuda", dtype=torch.float16), xp, fp)
