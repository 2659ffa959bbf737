"""Proposal-guided rendering: a small MLP predicts where along each ray the
compositing weights of the plain renderer lie, and 32 fine SDF samples are
drawn there instead of the plain renderer's uniform ~0.55 N.

Pieces (all HIP kernels behind the C-ABI):
  ray_state()       RayCut of every ray (origin, direction, in-cube count)
  plain_bins()      plain fused forward + per-ray 64-bin compositing weights
                    (the training targets)
  Proposal          float32 master parameters, bf16 MFMA inference
                    (logits at the 64 bin centres), hand-written backward
                    and Adam, trained on the GPU
  sample()          inverse-CDF fine sample positions t [n_rays, K]
  ListedRender      fused forward / recompute adjoint on explicit samples

The plain uniform renderer stays the fallback and the reference the
proposal's PSNR is judged against.
"""
from __future__ import annotations

import math
from typing import Any

import torch
from torch import autograd

from . import _lib
from . import fused as _fused
from .fused import SCATTER_DET, SCATTER_TILE, backward_det, camera_buffer, _timed
from .qsdvr import _ptr, _stream, _t

NBINS = 64


def _n_rays(n_cams, W, H, ray_pixel):
    if ray_pixel is None:
        return n_cams * W * H
    return _t("ray_pixel", ray_pixel, torch.int32, ndim=1).shape[0]


def _check_state(rs, n=None):
    """The ray state is [n, 8] float32 (origin, count bits, direction, 0)."""
    return _t("ray state", rs, torch.float32, ndim=2, shape=(n, 8)).shape[0]


def _check_grids(sdf_grid, data_grid=None):
    _t("sdf_grid", sdf_grid, torch.float32, ndim=4)
    R = sdf_grid.shape[0]
    if tuple(sdf_grid.shape) != (R, R, R, 7):
        raise ValueError(f"sdf_grid must be [R,R,R,7], got {tuple(sdf_grid.shape)}")
    if data_grid is not None:
        _t("data_grid", data_grid, torch.float32, ndim=2, shape=((R + 1) ** 3, None))
    return R


def _check_list(tlist, n):
    """Fine sample positions [n, K] float32 (K >= 2), one row per ray."""
    if tlist is None:
        return 0
    _t("tlist", tlist, torch.float32, ndim=2, shape=(n, None))
    if tlist.shape[1] < 2:
        raise ValueError("tlist needs at least 2 samples per ray")
    return int(tlist.shape[1])


def ray_state(cam_buf, n_cams, W, H, interval, ray_pixel=None):
    """[n_rays, 8] float32: origin, count bits, direction, 0."""
    n = _n_rays(n_cams, W, H, ray_pixel)
    out = torch.empty((n, 8), dtype=torch.float32, device=cam_buf.device)
    _lib.call("qsdvr_ray_state", _ptr(cam_buf), n_cams, W, H, _ptr(ray_pixel), n,
              float(interval), _ptr(out), _stream())
    return out


def ray_counts(rs):
    """In-cube step counts (uint32 bits stored in column 3)."""
    return rs[:, 3].contiguous().view(torch.int32)


def plain_bins(sdf_grid, data_grid, cam_buf, n_cams, W, H, interval, s, ray_pixel=None):
    """(colour [n,3], depth [n], bins [n,64]) of the plain uniform renderer."""
    n = _n_rays(n_cams, W, H, ray_pixel)
    _check_grids(sdf_grid, data_grid)
    dev = sdf_grid.device
    color = torch.empty((n, 3), dtype=torch.float32, device=dev)
    depth = torch.empty((n,), dtype=torch.float32, device=dev)
    bins = torch.zeros((n, NBINS), dtype=torch.float32, device=dev)
    _lib.call("qsdvr_fused_forward_bins", _ptr(cam_buf), n_cams, W, H, _ptr(ray_pixel), n,
              _ptr(sdf_grid), _ptr(data_grid), sdf_grid.shape[0], data_grid.shape[-1],
              float(interval), float(s), _ptr(color), _ptr(depth), _ptr(bins), _stream())
    return color, depth, bins


def plain_bins_sdf(sdf_grid, rs, interval, s):
    """bins [n, 64] of the plain uniform renderer's compositing weights for
    the rays of the ray state ``rs``, from the SDF alone (the training
    targets; equal to ``plain_bins``'s to float32 rounding)."""
    n = _check_state(rs)
    _check_grids(sdf_grid)
    bins = torch.empty((n, NBINS), dtype=torch.float32, device=rs.device)
    with _timed("plain_bins"):
        _lib.call("qsdvr_plain_bins", _ptr(rs), n, _ptr(sdf_grid), sdf_grid.shape[0],
                  float(interval), float(s), _ptr(bins), _stream())
    return bins


class Proposal:
    """The proposal MLP: float32 master parameters on the device, a packed
    bf16 copy for the MFMA kernels, Adam moments."""

    def __init__(self, device="cuda", seed=0, theta=None):
        self.P = int(_lib.fn("qsdvr_proposal_param_count")())
        if theta is None:
            theta = init_params(self.P, seed)
        self.theta = theta.to(device=device, dtype=torch.float32).contiguous()
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.step = 0
        nb = int(_lib.fn("qsdvr_proposal_packed_bytes")())
        self.packed = torch.empty(nb, dtype=torch.uint8, device=device)
        self.pack()

    def pack(self):
        _lib.call("qsdvr_proposal_pack", _ptr(self.theta), _ptr(self.packed), _stream())

    def logits(self, rs, interval, train_buffers=None):
        """Logits [n_rays, 64] at the bin centres of the rays in ``rs``."""
        n = _check_state(rs)
        z = torch.empty((n, NBINS), dtype=torch.float32, device=rs.device)
        acts = train_buffers if train_buffers is not None else (None,) * 4
        with _timed("proposal_mlp"):
            _lib.call("qsdvr_proposal_forward", _ptr(rs), n, float(interval), _ptr(self.packed),
                      _ptr(z), *[_ptr(a) for a in acts], _stream())
        return z

    def grad(self, rs, interval, target):
        """(loss [n], grad [P], logits) of the proposal loss against the bin
        targets ``target`` [n, 64]: per ray the log-likelihood of the plain
        renderer's bin weights under the composited proposal weights
        w_b = T_b (1 - exp(-softplus(z_b))) plus the pass-through outcome,
        through the hand-written forward and backward kernels."""
        n = _check_state(rs)
        dev = rs.device
        acts = [torch.empty((n * NBINS, 64), dtype=torch.float32, device=dev) for _ in range(4)]
        z = self.logits(rs, interval, acts)
        ws_n = int(_lib.fn("qsdvr_proposal_workspace_size")(n))
        ws = torch.empty(ws_n, dtype=torch.uint8, device=dev)
        g = torch.empty(self.P, dtype=torch.float32, device=dev)
        loss = torch.empty(n, dtype=torch.float32, device=dev)
        _t("target", target, torch.float32, shape=(n, NBINS))
        _lib.call("qsdvr_proposal_backward", _ptr(z), _ptr(target), n, _ptr(self.theta),
                  *[_ptr(a) for a in acts], _ptr(ws), ws_n, _ptr(g), _ptr(loss), _stream())
        return loss, g, z

    def adam(self, g, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        _lib.call("qsdvr_adam_step", _ptr(self.theta), _ptr(g), _ptr(self.m), _ptr(self.v),
                  self.P, float(lr), float(b1), float(b2), float(eps), self.step, _stream())
        self.pack()

    def state_dict(self):
        return {"theta": self.theta.detach().cpu(), "m": self.m.cpu(), "v": self.v.cpu(),
                "step": self.step}

    def load_state_dict(self, sd):
        self.theta.copy_(sd["theta"])
        self.m.copy_(sd["m"])
        self.v.copy_(sd["v"])
        self.step = int(sd["step"])
        self.pack()


def init_params(P, seed=0):
    """He-uniform weights, zero biases (W1 fan-in 39, the others 64)."""
    g = torch.Generator().manual_seed(seed)
    shapes = [(64, 39), (64,), (64, 64), (64,), (64, 64), (64,), (64,), (1,)]
    parts = []
    for i, shp in enumerate(shapes):
        if len(shp) == 1 and i != 6:
            parts.append(torch.zeros(shp))
        else:
            fan = shp[-1] if len(shp) == 2 else 64
            parts.append((torch.rand(shp, generator=g) * 2 - 1) * math.sqrt(6.0 / fan))
    theta = torch.cat([p.reshape(-1) for p in parts])
    assert theta.numel() == P
    return theta


# Sampler defaults: no uniform mixture (eps = 0 beat eps = 0.05 by 0.1 dB at
# 384^3 and by 0.3-0.8 dB at 64^3-128^3) and the weights unsharpened.
# gamma > 1 concentrates the samples where the network's bins peak: at
# 384^3 / s = 40 / 48 samples gamma = 2 turns a 0.2 dB PSNR loss against the
# 384-sample plain render into a 0.1 dB gain, but at 64^3 it costs 3 dB,
# so it stays a per-scene choice.
SAMPLE_EPS = 0.0
SAMPLE_GAMMA = 1.0


def sample(z, rs, interval, K=32, eps=SAMPLE_EPS, gamma=SAMPLE_GAMMA):
    """Fine sample positions t [n_rays, K] (increasing, t0 = 0, t_{K-1} = L)
    from the pdf (1 - eps) w^gamma / sum w^gamma + eps / 64."""
    n = _check_state(rs)
    _t("logits", z, torch.float32, ndim=2, shape=(n, NBINS))
    if K < 3:
        raise ValueError("the sampler draws at least 3 samples per ray")
    t = torch.empty((n, K), dtype=torch.float32, device=rs.device)
    with _timed("proposal_sample"):
        _lib.call("qsdvr_proposal_sample", _ptr(z), _ptr(rs), n, float(interval), int(K),
                  float(eps), float(gamma), _ptr(t), _stream())
    return t


class ListedRender(autograd.Function):
    """Fused render on explicit per-ray sample positions (``tlist`` [n, K])
    or, with tlist None, on the uniform samples of the ray state."""

    @staticmethod
    def forward(ctx: Any, sdf_grid, data_grid, cam_buf, n_cams, W, H, ray_pixel, rs, tlist,
                interval, s, scatter_mode=SCATTER_TILE):
        n = _check_state(rs)
        if _n_rays(n_cams, W, H, ray_pixel) != n:
            raise ValueError("the ray state must hold one row per ray")
        K = _check_list(tlist, n)
        _check_grids(sdf_grid, data_grid)
        dev = sdf_grid.device
        color = torch.empty((n, 3), dtype=torch.float32, device=dev)
        depth = torch.empty((n,), dtype=torch.float32, device=dev)
        cost = None
        with _timed("listed_forward"):
            if _fused._ordered(ray_pixel, scatter_mode, data_grid.shape[-1]):
                # whole frames: record the tiles' costs for the ordered adjoint
                cost = torch.empty(_fused._tile_count(n_cams, W, H), dtype=torch.int32,
                                   device=dev)
                _lib.call("qsdvr_fused_forward_costed", _ptr(cam_buf), n_cams, W, H, _ptr(rs),
                          _ptr(tlist), K, _ptr(sdf_grid), _ptr(data_grid), sdf_grid.shape[0],
                          data_grid.shape[-1], float(interval), float(s), _ptr(color),
                          _ptr(depth), _ptr(cost), _stream())
            else:
                _lib.call("qsdvr_fused_forward_listed", _ptr(cam_buf), n_cams, W, H,
                          _ptr(ray_pixel), n, _ptr(rs), _ptr(tlist), K, _ptr(sdf_grid),
                          _ptr(data_grid), sdf_grid.shape[0], data_grid.shape[-1],
                          float(interval), float(s), _ptr(color), _ptr(depth), _stream())
        ctx.save_for_backward(sdf_grid, data_grid, cam_buf, ray_pixel, rs, tlist, color, depth,
                              cost)
        ctx.args = (n_cams, W, H, n, K, float(interval), float(s), int(scatter_mode))
        ctx.set_materialize_grads(False)
        return color, depth

    @staticmethod
    def backward(ctx: Any, gcolor, gdepth):
        sdf_grid, data_grid, cam_buf, ray_pixel, rs, tlist, color, depth, cost = ctx.saved_tensors
        n_cams, W, H, n, K, interval, s, mode = ctx.args
        none = (None,) * 10
        if gcolor is None and gdepth is None:
            return (None, None) + none
        gq = torch.zeros_like(sdf_grid)
        gd = torch.zeros_like(data_grid)
        if gcolor is None:
            gcolor = torch.zeros((n, 3), dtype=torch.float32, device=sdf_grid.device)
        _t("colour gradient", gcolor.contiguous(), torch.float32, shape=(n, 3))
        gdep = gdepth.contiguous() if gdepth is not None else None
        if gdep is not None:
            _t("depth gradient", gdep, torch.float32, ndim=1, shape=(n,))
        if mode == SCATTER_DET:
            with _timed("listed_backward"):
                backward_det(cam_buf, n_cams, W, H, ray_pixel, rs, tlist, sdf_grid, data_grid,
                             interval, s, color, depth, gcolor.contiguous(), gdep, gq, gd)
            return (gq, gd) + none
        with _timed("listed_backward"):
            if cost is not None:
                ws, nb = _fused._order_workspace(n_cams, W, H, sdf_grid.device)
                _lib.call("qsdvr_fused_backward_ordered", _ptr(cam_buf), n_cams, W, H, _ptr(rs),
                          _ptr(tlist), K, _ptr(sdf_grid), _ptr(data_grid), sdf_grid.shape[0],
                          data_grid.shape[-1], interval, s, _ptr(color), _ptr(depth),
                          _ptr(gcolor.contiguous()), _ptr(gdep), _ptr(gq), _ptr(gd), _ptr(cost),
                          _ptr(ws), nb, _stream())
            else:
                _lib.call("qsdvr_fused_backward_listed", _ptr(cam_buf), n_cams, W, H,
                          _ptr(ray_pixel), n, _ptr(rs), _ptr(tlist), K, _ptr(sdf_grid),
                          _ptr(data_grid), sdf_grid.shape[0], data_grid.shape[-1], interval, s,
                          _ptr(color), _ptr(depth), _ptr(gcolor.contiguous()), _ptr(gdep),
                          _ptr(gq), _ptr(gd), mode, _stream())
        return (gq, gd) + none


def render_proposal(sdf_grid, data_grid, cam_buf, n_cams, W, H, interval, s, proposal,
                    K=32, eps=SAMPLE_EPS, ray_pixel=None, scatter_mode=SCATTER_TILE,
                    gamma=SAMPLE_GAMMA):
    """Proposal-guided render: ray state -> MLP logits -> fine samples ->
    listed fused render.  Returns (colour, depth); differentiable in the
    grids."""
    rs = ray_state(cam_buf, n_cams, W, H, interval, ray_pixel)
    z = proposal.logits(rs, interval)
    t = sample(z, rs, interval, K, eps, gamma)
    return ListedRender.apply(sdf_grid, data_grid, cam_buf, n_cams, W, H, ray_pixel, rs, t,
                              interval, s, scatter_mode)


def train(proposal, sdf_grid, data_grid, cams, interval, s, steps=4000, batch=8192, lr=1e-2,
          seed=0, log_every=0):
    """Train the proposal on the plain renderer's compositing weights of
    random rays of ``cams`` (targets built on the fly on the GPU).  Returns
    the per-step mean loss per unit weight."""
    dev = sdf_grid.device
    cam_buf = camera_buffer(cams, dev)
    W, H = int(cams[0].width), int(cams[0].height)
    n_all = len(cams) * W * H
    g = torch.Generator(device=dev).manual_seed(seed)
    hist = []
    for it in range(steps):
        pick = torch.randint(0, n_all, (batch,), generator=g, device=dev, dtype=torch.int64)
        pick = pick.to(torch.int32)
        rs = ray_state(cam_buf, len(cams), W, H, interval, pick)
        q = plain_bins_sdf(sdf_grid.detach(), rs, interval, s)
        loss, grad, _ = proposal.grad(rs, interval, q)
        proposal.adam(grad, lr=lr if it < int(0.8 * steps) else lr * 0.2)
        if log_every and (it % log_every == 0 or it == steps - 1):
            hist.append((it, float(loss.sum() / q.sum().clamp_min(1e-12))))
    return hist


def psnr(x, ref, data_range=None):
    """PSNR in dB with the reference's value range as the peak (the
    synthetic appearance is not limited to [0, 1])."""
    ref = ref.double()
    mse = ((x.double() - ref) ** 2).mean()
    rng = (ref.max() - ref.min()) if data_range is None else torch.as_tensor(data_range)
    return float(10 * torch.log10(rng ** 2 / mse))
