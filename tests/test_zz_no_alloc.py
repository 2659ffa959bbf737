"""The allocation-free RayCut (qsdvr_ray_cut_noalloc) against qsdvr_ray_cut,
through the C ABI: two calls in a row, the second on a smaller frame with a
mask, each bit-exact against the workspace-taking entry point on the same
camera; sizes beyond the static workspace are refused before any work."""
import ctypes as C

import numpy as np
import pytest
import torch

from qsdvr_amd import _lib, qsdvr, scenes

pytestmark = pytest.mark.gpu


def _outputs(P):
    return (torch.zeros((P, 4), dtype=torch.int32, device="cuda"),
            torch.zeros((P, 3), dtype=torch.float32, device="cuda"),
            torch.zeros((P, 3), dtype=torch.float32, device="cuda"))


def _noalloc(cam, mask, interval):
    lib = _lib.load()
    a = qsdvr.CameraArgs.from_camera(cam)
    P = a.width * a.height
    ray, ori, dirs = _outputs(P)
    m = None if mask is None else np.ascontiguousarray(mask.numpy().astype(np.uint8))
    st, rt = C.c_int64(-1), C.c_int64(-1)
    status = lib.qsdvr_ray_cut_noalloc(
        C.byref(a.to_struct()), C.c_void_p(0) if m is None else C.c_void_p(m.ctypes.data),
        C.c_void_p(ray.data_ptr()), C.c_void_p(ori.data_ptr()), C.c_void_p(dirs.data_ptr()),
        float(interval), P, C.byref(st), C.byref(rt),
        C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert status == 0, lib.qsdvr_last_error().decode()
    torch.cuda.synchronize()
    return ray.cpu(), ori.cpu(), dirs.cpu(), st.value, rt.value


def _with_workspace(cam, mask, interval):
    a = qsdvr.CameraArgs.from_camera(cam)
    P = a.width * a.height
    ray, ori, dirs = _outputs(P)
    qsdvr.RayCut(a, mask, ray, ori, dirs, interval)
    torch.cuda.synchronize()
    return ray.cpu(), ori.cpu(), dirs.cpu(), a.sdfCount, a.renderCount


def test_noalloc_ray_cut_matches_ray_cut_twice():
    interval = scenes.interval_for(384)
    big = scenes.orbit_cameras(50, 1920, 1080)[7]
    small = scenes.orbit_cameras(100, 640, 480)[31]
    mask = torch.rand(640 * 480, generator=torch.Generator().manual_seed(3)) > 0.3
    for cam, m in ((big, None), (small, mask)):
        ref = _with_workspace(cam, m, interval)
        got = _noalloc(cam, m, interval)
        assert got[3] == ref[3] and got[4] == ref[4] and ref[3] > 0
        for g, r in zip(got[:3], ref[:3]):
            assert torch.equal(g, r)


def test_noalloc_ray_cut_refuses_oversized_frames():
    lib = _lib.load()
    limit = int(lib.qsdvr_ray_cut_noalloc_max_pixels())
    assert limit == 1 << 23
    cam = scenes.orbit_cameras(1, 4097, 2048)[0]          # 8,390,656 > 2^23 pixels
    a = qsdvr.CameraArgs.from_camera(cam)
    st, rt = C.c_int64(-1), C.c_int64(-1)
    status = lib.qsdvr_ray_cut_noalloc(C.byref(a.to_struct()), None, None, None, None, 0.01,
                                       4097 * 2048, C.byref(st), C.byref(rt), None)
    assert _lib.STATUS_NAMES.get(status) == "ERANGE"
    assert "static workspace" in lib.qsdvr_last_error().decode()
    assert st.value == -1 and rt.value == -1
