"""The drop-in boundary on the CPU: library exports, argument scraping, error
types and messages of the reference's dose_kernels module
(DoseCUDA/dose_kernels/dosemodule.cu:14-101,189-197), beam-model loaders.
No compute call needs a GPU here; calls that reach the engine on a machine
without a GPU must raise RuntimeError("CUDA error: ...")."""
import os
import re

import numpy as np
import pytest

from dosecuda_amd import _lib, dose_kernels
from dosecuda_amd.beam_models import (PhotonBeamModel, ProtonBeamModel, density_from_hu,
                                      rlsp_from_hu, square_field_mlc, PhotonControlPoint)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "dosehip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    L = _lib.load()
    syms = _header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, s


def test_library_reports_version():
    assert _lib.load().dh_version() >= 10000


class _Vol:
    def __init__(self, shape=(4, 5, 6)):
        self.voxel_data = np.ones(shape, np.float32)
        self.origin = np.zeros(3, np.float32)
        self.spacing = np.full(3, 2.0, np.float32)


class _Beam:
    def __init__(self):
        self.iso = np.zeros(3, np.float32)
        self.gantry_angle = 0.0
        self.couch_angle = 0.0
        self.spot_list = np.array([[0, 0, 1, 10]], np.float32)


def test_raytrace_argument_errors(proton_model):
    v, b = _Vol(), _Beam()
    bad = _Vol()
    bad.voxel_data = bad.voxel_data.astype(np.float64)
    with pytest.raises(ValueError, match=r"^'voxel_data' must be 3-dimensional and of type float\.$"):
        dose_kernels.proton_raytrace_cuda(proton_model, bad, b, 0)
    bad = _Vol()
    bad.spacing = np.full((1, 3), 2.0, np.float32)
    with pytest.raises(ValueError, match="'spacing' must be 1-dimensional"):
        dose_kernels.proton_raytrace_cuda(proton_model, bad, b, 0)
    nb = _Beam()
    del nb.iso
    with pytest.raises(AttributeError):
        dose_kernels.proton_raytrace_cuda(proton_model, v, nb, 0)
    nb = _Beam()
    nb.gantry_angle = "90"
    with pytest.raises(TypeError):
        dose_kernels.proton_raytrace_cuda(proton_model, v, nb, 0)
    with pytest.raises(TypeError):
        dose_kernels.proton_raytrace_cuda(proton_model, v, b, 0.0)
    with pytest.raises(TypeError):
        dose_kernels.proton_raytrace_cuda(proton_model, v, b)


def test_spot_argument_errors(proton_model):
    v, b = _Vol(), _Beam()
    nb = _Beam()
    nb.spot_list = np.array([0, 0, 1, 10], np.float32)   # 1-D after a single addSingleSpot
    with pytest.raises(ValueError, match="'spot_list' must be 2-dimensional"):
        dose_kernels.proton_spot_cuda(proton_model, v, v, nb, 0)

    class M:
        VSADX, VSADY = 1340.0, 1930.0
        divergence_params = proton_model.divergence_params
        lut_depths = proton_model.lut_depths.astype(np.float64)
        lut_sigmas = proton_model.lut_sigmas
        lut_idds = proton_model.lut_idds
    with pytest.raises(ValueError, match="'lut_depths' must be 2-dimensional and of type float"):
        dose_kernels.proton_spot_cuda(M(), v, v, b, 0)


def test_photon_argument_errors(photon_model):
    v = _Vol()
    cp = PhotonControlPoint(np.zeros(3), 100.0, square_field_mlc(photon_model), 0.0, 0.0, 0.0)

    class M:
        pass
    m = M()
    for k, val in vars(photon_model).items():
        setattr(m, k, val)
    del m.mlc_transmission
    with pytest.raises(AttributeError):
        dose_kernels.photon_dose_cuda(m, v, cp, 0)
    m.mlc_transmission = 0.02
    m.kernel = np.asarray(m.kernel, np.float64)
    with pytest.raises(ValueError, match="'kernel' must be 2-dimensional and of type float"):
        dose_kernels.photon_dose_cuda(m, v, cp, 0)


def test_no_gpu_raises_cuda_error(proton_model):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match=r"^CUDA error: "):
        dose_kernels.proton_raytrace_cuda(proton_model, _Vol(), _Beam(), 0)


def test_proton_model_tables(proton_model):
    m = proton_model
    assert (m.VSADX, m.VSADY) == (1340.0, 1930.0)
    assert m.lut_depths.shape == (98, 400) and m.lut_depths.dtype == np.float32
    assert m.divergence_params.shape == (98, 5)
    np.testing.assert_array_equal(m.lut_depths[:, 0], 0.0)
    np.testing.assert_array_equal(m.lut_depths[:, 399], 399.0)
    assert m.energy_labels[0] == pytest.approx(70.2) and m.energyIDFromLabel(101.3) == 34
    assert m.divergence_params[34, 1] == pytest.approx(78.32)


def test_photon_model_tables(photon_model):
    m = photon_model
    assert m.kernel.shape == (6, 6) and m.kernel.flags.f_contiguous
    # reference reads the kernel buffer column-major: theta column first
    np.testing.assert_allclose(m.kernel.ravel(order="K")[:6],
                               [1.875, 20.625, 43.125, 61.875, 88.125, 106.875])
    assert m.n_mlc_pairs == 60 and m.mlc_distance == 491.0
    assert m.mu_calibration == pytest.approx(4.143232e-3)
    assert m.has_xjaws is True


def test_square_field_mlc_and_output_factor(photon_model):
    mlc = square_field_mlc(photon_model, 10.0, 10.0)
    assert mlc.shape == (60, 4) and mlc.flags.f_contiguous
    open_ = mlc[:, 1] > mlc[:, 0]
    # offsets within +-50 mm are open: 20 pairs of 5 mm
    assert open_.sum() == 20 and np.all(mlc[open_, 3] == 5.0)
    cp = PhotonControlPoint(np.zeros(3), 100.0, mlc, 0.0, 0.0, 0.0)
    # 100 x 100 mm: equivalent square 4A/P = 100 -> tabulated value
    assert photon_model.outputFactor(cp) == pytest.approx(0.979229534, rel=1e-6)


def test_hu_conversion_clamps():
    hu = np.array([-2000.0, -1000.0, 0.0, 1000.0, 5000.0])
    r = rlsp_from_hu(hu)
    assert r.dtype == np.float32
    np.testing.assert_allclose(r[[0, 1, 2, 4]], [0.00113, 0.00113, 1.00653, 2.36825], rtol=1e-6)
    np.testing.assert_allclose(density_from_hu(hu)[2], 1.00653, rtol=1e-6)


def test_influence_count_many_argument_errors():
    """The multi-field count rejects bad arguments before touching a device
    (so the checks run here without a GPU) and does nothing for no fields."""
    L = _lib.load()
    assert L.dh_influence_count_many(None, 0, None, 1.0, None, 0, None, None, None) == 0
    assert L.dh_influence_count_many(None, -1, None, 1.0, None, 0, None, None, None) != 0
    assert "field count" in _lib.last_error()
    assert L.dh_influence_count_many(None, 2, None, 1.0, None, 0, None, None, None) != 0
    assert "null argument" in _lib.last_error()
    import ctypes as C
    fields = (C.c_void_p * 1)(None)
    wets = (C.c_void_p * 1)(None)
    nnz = np.zeros(1, np.int64)
    dummy = np.zeros(4, np.float32)
    rc = L.dh_influence_count_many(fields, 1, _lib.ptr(dummy), 1.0, wets, 0, _lib.ptr(nnz), None, None)
    assert rc != 0 and "null field or WET buffer" in _lib.last_error()
    rc = L.dh_influence_count_many(fields, 1, _lib.ptr(dummy), -1.0, wets, 0, _lib.ptr(nnz), None, None)
    assert rc != 0 and "margin_scale" in _lib.last_error()
