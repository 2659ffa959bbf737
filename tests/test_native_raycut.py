"""The closed-form ray cut of the HIP library (qsdvr_amd/csrc/raygen.h:
cut_closed_form, float64 slab bounds + the exact float32 predicate near the
ends) against the reference's walk over every candidate step
(kernel/camera.cu:43-54, restated as cut_walk in the same header), compiled
for the host: counts and first in-cube steps must be bit-identical on
random, axis-aligned, corner-crossing and face-grazing rays."""
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
