"""The C-ABI library loads and exports every entry point declared in
include/qsdvr_hip.h; argument validation answers without touching a GPU."""
import ctypes as C
import re
from pathlib import Path

import pytest

from qsdvr_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "qsdvr_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsdvr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_op_surface():
    names = declared_functions()
    for n in ("qsdvr_ray_cut", "qsdvr_generate_ray_points", "qsdvr_layer_to_grid_forward",
              "qsdvr_layer_to_grid_backward", "qsdvr_sample_sdf_forward",
              "qsdvr_sample_sdf_backward", "qsdvr_sample_normal_forward",
              "qsdvr_sample_normal_backward", "qsdvr_grid_interpolation_forward",
              "qsdvr_grid_interpolation_backward", "qsdvr_shader_forward",
              "qsdvr_shader_backward", "qsdvr_ray_aggregate_forward",
              "qsdvr_ray_aggregate_backward"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows every one of them
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_version_and_validation_without_gpu():
    lib = _lib.load()
    assert lib.qsdvr_abi_version() == 1
    st = lib.qsdvr_sample_sdf_forward(None, None, None, None, -1, 8, 2.0, None)
    assert st == -1 and b"bad sizes" in lib.qsdvr_last_error()
    st = lib.qsdvr_shader_forward(None, None, None, None, None, 10, 31, None)
    assert st == -1 and b"channels" in lib.qsdvr_last_error()
    st = lib.qsdvr_generate_ray_points(None, None, None, None, None, None, None, None, 5, 0, 0,
                                       0.1, 0, None)
    assert st == -1
    assert lib.qsdvr_ray_cut_workspace_size(1000) >= 1000
    assert lib.qsdvr_layer_to_grid_workspace_size(128) >= 6 * 128 * 128 * 4
    # zero-size work is a no-op success
    assert lib.qsdvr_sample_sdf_forward(None, None, None, None, 0, 8, 2.0, None) == 0


def test_fused_rejects_bad_arguments():
    lib = _lib.load()
    st = lib.qsdvr_fused_forward(None, 1, 8, 8, None, 64, None, None, 4, 32, 0.1, 5.0, None,
                                 None, None)
    assert st == -1 and b"null" in lib.qsdvr_last_error()
    dummy = C.c_void_p(16)
    st = lib.qsdvr_fused_forward(dummy, 1, 8, 8, None, 65, dummy, dummy, 4, 32, 0.1, 5.0,
                                 dummy, None, None)
    assert st == -1 and b"n_rays" in lib.qsdvr_last_error()
    st = lib.qsdvr_fused_forward(dummy, 1, 8, 8, None, 64, dummy, dummy, 4, 33, 0.1, 5.0,
                                 dummy, None, None)
    assert st == -1 and b"channels" in lib.qsdvr_last_error()


def test_deterministic_backward_validates_on_the_host():
    lib = _lib.load()
    small = lib.qsdvr_fused_det_workspace_size(1000, 10_000, 32, 64)
    large = lib.qsdvr_fused_det_workspace_size(1000, 100_000, 32, 64)
    # records alone: 2 quadric (4 + 32 B) and one appearance (4 + 144 B) per slot
    assert small >= 10_000 * (2 * 36 + 148) and large > small
    assert lib.qsdvr_fused_det_workspace_size(1000, 100, 32, 513) == 0
    dummy = C.c_void_p(16)
    args = [dummy, 1, 8, 8, None, 64]
    st = lib.qsdvr_fused_backward_det(*args, None, None, 0, dummy, dummy, 4, 32, 0.1, 5.0,
                                      dummy, None, dummy, None, dummy, dummy, 100, dummy, 1 << 20,
                                      None)
    assert st == -1 and b"ray state" in lib.qsdvr_last_error()
    st = lib.qsdvr_fused_backward_det(*args, dummy, None, 0, dummy, dummy, 4, 32, 0.1, 5.0,
                                      dummy, None, dummy, None, dummy, dummy, 100, dummy, 64,
                                      None)
    assert st == -4 and b"too small" in lib.qsdvr_last_error()
    st = lib.qsdvr_fused_backward_det(*args, dummy, None, 0, dummy, dummy, 600, 32, 0.1, 5.0,
                                      dummy, None, dummy, None, dummy, dummy, 100, dummy, 1 << 30,
                                      None)
    assert st == -1 and b"512" in lib.qsdvr_last_error()


def test_sorted_scatters_validate_on_the_host():
    lib = _lib.load()
    nb = lib.qsdvr_sorted_scatter_workspace_size(1000, 64 ** 3)
    assert nb >= 4 * 4 * 1000
    assert lib.qsdvr_sorted_scatter_workspace_size(1000, 0) == 0
    dummy = C.c_void_p(16)
    # odd cell resolution and unsupported channel counts are refused
    st = lib.qsdvr_grid_interpolation_backward_sorted(dummy, dummy, dummy, dummy, 100, 10, 32,
                                                      dummy, 1 << 20, None)
    assert st == -1 and b"even" in lib.qsdvr_last_error()
    st = lib.qsdvr_grid_interpolation_backward_sorted(dummy, dummy, dummy, dummy, 100, 9, 31,
                                                      dummy, 1 << 20, None)
    assert st == -1
    st = lib.qsdvr_sample_sdf_backward_sorted(dummy, dummy, dummy, dummy, dummy, 100, 512, 8.0,
                                              dummy, 16, None)
    assert st == -4 and b"too small" in lib.qsdvr_last_error()
    assert lib.qsdvr_sample_normal_backward_sorted(None, None, None, None, None, 0, 512, None, 0,
                                                   None) == 0
    # the presorted SampleNormalBackward: an even resolution, its workspace
    # size checked against GridInterpolationBackward's, nothing to do for n = 0
    assert lib.qsdvr_sample_normal_backward_presorted(None, None, None, 0, 8, None, 0, None) == 0
    st = lib.qsdvr_sample_normal_backward_presorted(dummy, dummy, dummy, 100, 9, dummy, 1 << 20,
                                                    None)
    assert st == -1 and b"even" in lib.qsdvr_last_error()
    st = lib.qsdvr_sample_normal_backward_presorted(dummy, dummy, dummy, 100, 8, dummy, 16, None)
    assert st == -4 and b"too small" in lib.qsdvr_last_error()
    st = lib.qsdvr_sample_normal_backward_presorted(None, dummy, dummy, 100, 8, dummy, 1 << 20,
                                                    None)
    assert st == -1 and b"null" in lib.qsdvr_last_error()
    # the sorted GridInterpolationForward (32 channels, even resolution, 16-B
    # aligned grid and output) and the backward on its sort
    fwd = lib.qsdvr_grid_interpolation_forward_sorted
    assert fwd(None, None, None, None, 0, 9, 32, None, 0, None) == 0
    st = fwd(dummy, dummy, dummy, dummy, 100, 10, 32, dummy, 1 << 20, None)
    assert st == -1 and b"even" in lib.qsdvr_last_error()
    assert fwd(dummy, dummy, dummy, dummy, 100, 9, 53, dummy, 1 << 20, None) == -1
    st = fwd(C.c_void_p(20), dummy, dummy, dummy, 100, 9, 32, dummy, 1 << 30, None)
    assert st == -1 and b"aligned" in lib.qsdvr_last_error()
    st = fwd(dummy, dummy, dummy, dummy, 100, 9, 32, dummy, 16, None)
    assert st == -4 and b"too small" in lib.qsdvr_last_error()
    st = fwd(dummy, None, dummy, dummy, 100, 9, 32, dummy, 1 << 30, None)
    assert st == -1 and b"null" in lib.qsdvr_last_error()
    pre = lib.qsdvr_grid_interpolation_backward_presorted
    assert pre(None, None, 0, 9, 32, None, 0, None) == 0
    st = pre(dummy, dummy, 100, 10, 32, dummy, 1 << 20, None)
    assert st == -1 and b"even" in lib.qsdvr_last_error()
    st = pre(dummy, dummy, 100, 9, 32, dummy, 16, None)
    assert st == -4 and b"too small" in lib.qsdvr_last_error()
    st = pre(dummy, None, 100, 9, 53, dummy, 1 << 30, None)
    assert st == -1 and b"null" in lib.qsdvr_last_error()
