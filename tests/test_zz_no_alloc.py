"""The no-allocation proton WET entry point against the allocating one.

dh_proton_raytrace_dev_noalloc keeps its padded volume and unsmoothed WET in
the library's fixed device array; it must give dh_proton_raytrace_dev's WET
bit for bit, on two calls in a row (the second on a smaller volume, so any
stale contents of the shared array from the first call would show), and
refuse a volume beyond the array's size with DH_E_INVALID before any work.
Both entry points are called through the C ABI.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32
VSADX, VSADY = 1340.0, 1930.0


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _phantom(shape, s, seed):
    rng = np.random.default_rng(seed)
    rho = np.full(shape, 1.00653, F)
    k0, j0, i0 = (rng.integers(n // 4, n // 2) for n in shape)
    rho[k0:k0 + shape[0] // 3, j0:j0 + shape[1] // 5, i0:i0 + shape[2] // 3] = 1.5562
    rho[:, :, : shape[2] // 6] *= F(0.3)
    return rho


def _wet_pair(L, t, shape, s, ga, ta, seed):
    from dosecuda_amd import _lib
    rho = _phantom(shape, s, seed)
    iso = np.array([(shape[2] - 1) * s / 2 + 1.1, (shape[1] - 1) * s / 2 - 0.4,
                    (shape[0] - 1) * s / 2 + 0.3], F)
    g = _lib.grid(shape, s, (1.0, -2.0, 0.5))
    b = _lib.proton_beam(iso + F([1.0, -2.0, 0.5]), ga, ta, VSADX, VSADY)
    d_rho = t.from_numpy(rho).cuda()
    ref = t.full(shape, -7.0, dtype=t.float32, device="cuda")
    got = t.full(shape, 7.0, dtype=t.float32, device="cuda")
    stream = C.c_void_p(t.cuda.current_stream().cuda_stream)
    _lib.check(L.dh_proton_raytrace_dev(0, C.byref(g), C.byref(b), _lib.ptr(d_rho), _lib.ptr(ref),
                                        None, stream))
    _lib.check(L.dh_proton_raytrace_dev_noalloc(0, C.byref(g), C.byref(b), _lib.ptr(d_rho),
                                                _lib.ptr(got), stream))
    t.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_wet_equals_allocating_entry_point():
    t = _torch()
    from dosecuda_amd import _lib
    L = _lib.load()
    _lib.check(L.dh_init(0))
    cases = [((100, 104, 96), 2.5, 30.0, 10.0, 1), ((36, 44, 52), 2.0, 250.0, 340.0, 2)]
    for shape, s, ga, ta, seed in cases:
        print(f"[case] noalloc WET shape={shape} s={s} ga={ga} ta={ta}", flush=True)
        ref, got = _wet_pair(L, t, shape, s, ga, ta, seed)
        assert np.all(np.isfinite(ref)) and ref.max() > 1.0
        assert np.array_equal(ref.view(np.uint32), got.view(np.uint32)), shape


def test_noalloc_wet_rejects_oversized_volume():
    t = _torch()
    from dosecuda_amd import _lib
    L = _lib.load()
    shape = (160, 160, 160)          # 163^3 * 16 + 160^3 * 4 bytes > 64 MiB
    g = _lib.grid(shape, 2.0, (0.0, 0.0, 0.0))
    b = _lib.proton_beam((160.0, 160.0, 160.0), 0.0, 0.0, VSADX, VSADY)
    d = t.zeros(8, dtype=t.float32, device="cuda")   # never touched: refused before any launch
    rc = L.dh_proton_raytrace_dev_noalloc(0, C.byref(g), C.byref(b), _lib.ptr(d), _lib.ptr(d),
                                          C.c_void_p(t.cuda.current_stream().cuda_stream))
    assert rc == _lib.DH_E_INVALID
    assert "64 MiB" in _lib.last_error()
