"""Host builds of header-only device logic.

The closed-form ray cut of the HIP library (qsdvr_amd/csrc/raygen.h:
cut_closed_form, float64 slab bounds + the exact float32 predicate near the
ends) against the reference's walk over every candidate step
(kernel/camera.cu:43-54, restated as cut_walk in the same header): counts
and first in-cube steps must be bit-identical on random, axis-aligned,
corner-crossing and face-grazing rays.  And GridInterpolationBackward's sort
keys (qsdvr_amd/csrc/det_scatter.h: ClassKeys)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_closed_form_ray_cut_equals_walk(tmp_path):
    exe = tmp_path / "raycut_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT}/include",
                    f"-I{ROOT}/qsdvr_amd/csrc", f"{ROOT}/tests/native/raycut_check.cpp",
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe), "1000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_interpolation_sort_keys_invert(tmp_path):
    """GridInterpolationBackward's (parity class, cell) sort keys
    (qsdvr_amd/csrc/det_scatter.h: ClassKeys), compiled for the host: a
    permutation of [0, R^3) that decodes back to the cell, class-contiguous
    and rising with the cell index within a class."""
    exe = tmp_path / "classkeys_check"
    subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{ROOT}/include",
                    f"-I{ROOT}/qsdvr_amd/csrc", f"{ROOT}/tests/native/classkeys_check.cpp",
                    "-o", str(exe)], check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
