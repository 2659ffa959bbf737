"""Fused plain renderer: one forward kernel and one recompute-adjoint kernel
behind a ``torch.autograd.Function``.

Computes the same thing as the op-surface composition of models/render.py:
225-234 preceded by RayCut + GenerateRayPoints (kernel/camera.cu), for every
pixel of a batch of cameras (or for an explicit list of rays), but keeps no
per-sample tensors: the forward saves the per-ray colour and depth only and
the backward rebuilds alpha and T front to back from them.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Sequence

import numpy as np
import torch
from torch import autograd

from . import _lib
from .qsdvr import CameraArgs, _ptr, _stream, _t

SCATTER_TILE = 0     # tile-reduced atomics
SCATTER_ATOMIC = 1   # one atomic per sample and channel
SCATTER_DET = 2      # sorted-segment sums, bitwise reproducible

# Whole-frame tile-mode renders record each 8x8 tile's expected backward cost
# in the forward, and the adjoint then runs the tiles heaviest first
# (qsdvr_fused_forward_costed / qsdvr_fused_backward_ordered): a fixed launch
# order leaves the heaviest tiles to finish last.  TILE_ORDER False: the
# launch order; TILE_ORDER_CHANNELS: the appearance channel counts it is
# used for.  Not 53: the costed 53-channel forward compiles to a schedule
# that measured 3.3 ms slower on a 384^3 grid at 800x800 with 48 listed
# samples per ray, more than the ordering saves in its adjoint (0.7 ms).
TILE_ORDER = True
TILE_ORDER_CHANNELS = (32,)


def _ordered(ray_pixel, scatter_mode, channels):
    return (TILE_ORDER and ray_pixel is None and scatter_mode == SCATTER_TILE
            and channels in TILE_ORDER_CHANNELS)

# When set to a list, every fused kernel launch appends
# (name, start_event, end_event) recorded on the launching stream.
TIMERS = None


def _tile_count(n_cams, width, height):
    return n_cams * ((width + 7) // 8) * ((height + 7) // 8)


def _order_workspace(n_cams, width, height, device):
    nb = int(_lib.fn("qsdvr_fused_order_workspace_size")(n_cams, width, height))
    return torch.empty(max(nb, 1), dtype=torch.uint8, device=device), nb


class _timed:
    def __init__(self, name):
        self.name = name

    def __enter__(self):
        if TIMERS is not None:
            self.a = torch.cuda.Event(enable_timing=True)
            self.b = torch.cuda.Event(enable_timing=True)
            self.a.record()

    def __exit__(self, *exc):
        if TIMERS is not None:
            self.b.record()
            TIMERS.append((self.name, self.a, self.b))


def camera_buffer(cams: Sequence, device="cuda") -> torch.Tensor:
    """Pack cameras (``scenes.Camera`` / ``CameraArgs``-like objects) into a
    device byte tensor holding a ``qsdvr_camera[]`` array."""
    n = len(cams)
    arr = (_lib.CameraStruct * n)()
    for i, c in enumerate(cams):
        a = c if isinstance(c, CameraArgs) else CameraArgs.from_camera(c)
        arr[i] = a.to_struct()
    raw = bytes(arr)
    return torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)


class FusedRender(autograd.Function):
    """color [n_rays,3], depth [n_rays] = render(sdf_grid, data_grid).

    Gradients flow to the quadric grid [R,R,R,7] and the appearance grid
    [(R+1)^3, C] (the reference's trainable tensors)."""

    @staticmethod
    def forward(ctx: Any, sdf_grid, data_grid, cam_buf, n_cams: int, width: int, height: int,
                ray_pixel, interval: float, s: float, scatter_mode: int = SCATTER_TILE):
        R = sdf_grid.shape[0]
        Cc = data_grid.shape[-1]
        _t("sdf_grid", sdf_grid, torch.float32)
        _t("data_grid", data_grid, torch.float32)
        if sdf_grid.numel() != R ** 3 * 7 or data_grid.numel() != (R + 1) ** 3 * Cc:
            raise ValueError("grid shapes must be [R,R,R,7] and [(R+1)^3, C]")
        if ray_pixel is None:
            n_rays = n_cams * width * height
        else:
            _t("ray_pixel", ray_pixel, torch.int32, ndim=1)
            n_rays = ray_pixel.shape[0]
        dev = sdf_grid.device
        color = torch.empty((n_rays, 3), dtype=torch.float32, device=dev)
        depth = torch.empty((n_rays,), dtype=torch.float32, device=dev)
        cost = None
        with _timed("fused_forward"):
            if _ordered(ray_pixel, scatter_mode, Cc):
                cost = torch.empty(_tile_count(n_cams, width, height), dtype=torch.int32,
                                   device=dev)
                _lib.call("qsdvr_fused_forward_costed", _ptr(cam_buf), n_cams, width, height,
                          None, None, 0, _ptr(sdf_grid), _ptr(data_grid), R, Cc, float(interval),
                          float(s), _ptr(color), _ptr(depth), _ptr(cost), _stream())
            else:
                _lib.call("qsdvr_fused_forward", _ptr(cam_buf), n_cams, width, height,
                          _ptr(ray_pixel), n_rays, _ptr(sdf_grid), _ptr(data_grid), R, Cc,
                          float(interval), float(s), _ptr(color), _ptr(depth), _stream())
        ctx.save_for_backward(sdf_grid, data_grid, cam_buf, ray_pixel, color, depth, cost)
        ctx.args = (n_cams, width, height, n_rays, R, Cc, float(interval), float(s),
                    int(scatter_mode))
        ctx.set_materialize_grads(False)
        return color, depth

    @staticmethod
    def backward(ctx: Any, gcolor, gdepth):
        sdf_grid, data_grid, cam_buf, ray_pixel, color, depth, cost = ctx.saved_tensors
        n_cams, width, height, n_rays, R, Cc, interval, s, mode = ctx.args
        dev = sdf_grid.device
        gq = torch.zeros_like(sdf_grid)
        gd = torch.zeros_like(data_grid)
        if gcolor is None and gdepth is None:
            return None, None, None, None, None, None, None, None, None, None
        if gcolor is None:
            gcolor = torch.zeros((n_rays, 3), dtype=torch.float32, device=dev)
        gcolor = gcolor.contiguous()
        gdep = gdepth.contiguous() if gdepth is not None else None
        if mode == SCATTER_DET:
            rs = torch.empty((n_rays, 8), dtype=torch.float32, device=dev)
            _lib.call("qsdvr_ray_state", _ptr(cam_buf), n_cams, width, height, _ptr(ray_pixel),
                      n_rays, float(interval), _ptr(rs), _stream())
            with _timed("fused_backward"):
                backward_det(cam_buf, n_cams, width, height, ray_pixel, rs, None, sdf_grid,
                             data_grid, interval, s, color, depth, gcolor, gdep, gq, gd)
            return gq, gd, None, None, None, None, None, None, None, None
        with _timed("fused_backward"):
            if cost is not None:
                ws, nb = _order_workspace(n_cams, width, height, dev)
                _lib.call("qsdvr_fused_backward_ordered", _ptr(cam_buf), n_cams, width, height,
                          None, None, 0, _ptr(sdf_grid), _ptr(data_grid), R, Cc, interval, s,
                          _ptr(color), _ptr(depth), _ptr(gcolor), _ptr(gdep), _ptr(gq), _ptr(gd),
                          _ptr(cost), _ptr(ws), nb, _stream())
            else:
                _lib.call("qsdvr_fused_backward", _ptr(cam_buf), n_cams, width, height,
                          _ptr(ray_pixel), n_rays, _ptr(sdf_grid), _ptr(data_grid), R, Cc,
                          interval, s, _ptr(color), _ptr(depth), _ptr(gcolor), _ptr(gdep),
                          _ptr(gq), _ptr(gd), mode, _stream())
        return gq, gd, None, None, None, None, None, None, None, None


def backward_det(cam_buf, n_cams, width, height, ray_pixel, rs, tlist, sdf_grid, data_grid,
                 interval, s, color, depth, gcolor, gdepth, gq, gd):
    """Deterministic adjoint (qsdvr_fused_backward_det) on the rays of the
    ray state ``rs``: sizes its workspace from the rays' sample counts (one
    host synchronisation) and adds into the zeroed ``gq`` and ``gd``."""
    n_rays = int(rs.shape[0])
    R = int(sdf_grid.shape[0])
    Cc = int(data_grid.shape[-1])
    K = int(tlist.shape[1]) if tlist is not None else 0
    cnt = rs[:, 3].contiguous().view(torch.int32).to(torch.int64)
    per_ray = torch.where(cnt >= 2, torch.full_like(cnt, K) if K else cnt, torch.zeros_like(cnt))
    cap = int(per_ray.sum())
    nbytes = int(_lib.fn("qsdvr_fused_det_workspace_size")(n_rays, cap, Cc, R))
    if nbytes == 0:
        raise ValueError(f"deterministic scatter: unsupported sizes (reso {R} must be <= 512)")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=sdf_grid.device)
    _lib.call("qsdvr_fused_backward_det", _ptr(cam_buf), n_cams, width, height, _ptr(ray_pixel),
              n_rays, _ptr(rs), _ptr(tlist), K, _ptr(sdf_grid), _ptr(data_grid), R, Cc,
              float(interval), float(s), _ptr(color), _ptr(depth), _ptr(gcolor), _ptr(gdepth),
              _ptr(gq), _ptr(gd), cap, _ptr(ws), nbytes, _stream())


def render(sdf_grid, data_grid, cams, interval: float, s: float, ray_pixel=None,
           scatter_mode: int = SCATTER_TILE, cam_buf=None):
    """Render every pixel of ``cams`` (list of cameras of one size) or the
    rays ``ray_pixel`` (int32 codes cam*W*H + pixel).  Returns (color, depth)
    with one row per ray, row-major per camera."""
    if cam_buf is None:
        cam_buf = camera_buffer(cams, sdf_grid.device)
    W, H = int(cams[0].width), int(cams[0].height)
    return FusedRender.apply(sdf_grid, data_grid, cam_buf, len(cams), W, H, ray_pixel,
                             interval, s, scatter_mode)
