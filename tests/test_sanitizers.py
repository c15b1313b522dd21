"""Sanitizer builds on the CPU (no GPU needed).

* The FP64 oracle (oracle/dose_oracle.c) built with AddressSanitizer and
  UndefinedBehaviorSanitizer and driven through every entry point on small
  odd-shaped grids (oracle/sanitize_main.c).
* The engine's GPU-free host set-up (dosecuda_amd/csrc/host_tables.cpp: beam
  geometry, spot sorting and layer import, TERMA arguments with the MLC
  rectangle merge, cone arguments, the influence matrix's table bounds)
  built with host-side AddressSanitizer + UndefinedBehaviorSanitizer and
  driven on seeded random inputs, including refused arguments
  (dosecuda_amd/csrc/sanitize_host.cpp).
Each must exit 0 with no sanitizer report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")


def _run(exe):
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out, out
    return out


def test_oracle_under_asan_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not found")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    assert "sanitize: ok" in _run(os.path.join(ROOT, "oracle", "_build", "sanitize_check"))


def test_host_setup_under_asan_ubsan():
    from dosecuda_amd import build as B
    try:
        B.hipcc()
    except RuntimeError:
        pytest.skip("hipcc not found")
    exe = B.build_sanitize_host()
    assert "sanitize host: ok" in _run(exe)
