"""GPU-resident training step of the plain renderer.

The reference trains with a host loop (test.py:27-37): its quadric
parameters are CPU tensors turned into the grid by LayerToGrid every step
(models/render.py:172-206), every step copies the ray lists host to device
(datasets/utils.py:16-40), and torch.optim.Adam updates the parameters.
Here one step stays on the device end to end:

  random ray batch (drawn on the device)
  -> [LayerToGrid on the device, for the reference's layer parameters]
  -> fused forward -> MSE -> recompute adjoint (any scatter mode)
  -> [LayerToGrid adjoint on the device]
  -> Adam (qsdvr_adam_step, HIP) on every parameter.

Parameters are either the quadric grid [R,R,R,7] itself or the reference's
separable layers (x, y, z [R], offset [4]), plus the appearance grid
[(R+1)^3, C].  Adam is the bias-corrected update of torch.optim.Adam.
"""
from __future__ import annotations

import torch

from . import _lib, fused
from .qsdvr import _ptr, _stream
from .render import LayerToGrid


class DeviceAdam:
    """Adam moments on the device and the HIP update kernel; the update is
    torch.optim.Adam's (bias-corrected, eps outside the square root)."""

    def __init__(self, params, lr=1e-2, betas=(0.9, 0.999), eps=1e-8):
        self.params = list(params)
        for p in self.params:
            if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous():
                raise ValueError("DeviceAdam: parameters must be contiguous float32 CUDA tensors")
        self.lr = float(lr)
        self.b1, self.b2 = map(float, betas)
        self.eps = float(eps)
        self.m = [torch.zeros_like(p) for p in self.params]
        self.v = [torch.zeros_like(p) for p in self.params]
        self.t = 0

    @torch.no_grad()
    def step(self):
        self.t += 1
        for p, m, v in zip(self.params, self.m, self.v):
            if p.grad is None:
                continue
            g = p.grad.contiguous()
            _lib.call("qsdvr_adam_step", _ptr(p), _ptr(g), _ptr(m), _ptr(v), p.numel(),
                      self.lr, self.b1, self.b2, self.eps, self.t, _stream())

    def zero_grad(self):
        for p in self.params:
            p.grad = None

    def state_dict(self):
        """Moments and step count (host copies), for checkpoints."""
        return {"t": self.t, "lr": self.lr, "betas": (self.b1, self.b2), "eps": self.eps,
                "m": [m.detach().cpu() for m in self.m], "v": [v.detach().cpu() for v in self.v]}

    def load_state_dict(self, sd):
        if len(sd["m"]) != len(self.m):
            raise ValueError("optimiser state has a different parameter list")
        self.t = int(sd["t"])
        self.lr = float(sd["lr"])
        self.b1, self.b2 = map(float, sd["betas"])
        self.eps = float(sd["eps"])
        for dst, src in zip(self.m + self.v, list(sd["m"]) + list(sd["v"])):
            if dst.shape != src.shape:
                raise ValueError("optimiser state shape mismatch")
            dst.copy_(src)


class FusedTrainer:
    """One training step = one random batch of rays over ``n_cams`` cameras
    (device buffer ``cam_buf``, all ``width`` x ``height``) against target
    colours ``target_fn(codes)`` (or explicit targets).

    ``sdf_grid``: trainable quadric grid, or None with ``layers`` = (x, y, z,
    offset) device tensors of the reference's parameterisation."""

    def __init__(self, data_grid, cam_buf, n_cams, width, height, interval, s,
                 sdf_grid=None, layers=None, lr=1e-2, scatter_mode=fused.SCATTER_DET,
                 seed=0):
        if (sdf_grid is None) == (layers is None):
            raise ValueError("give exactly one of sdf_grid and layers")
        self.data = data_grid.detach().requires_grad_(True)
        self.sdf = sdf_grid.detach().requires_grad_(True) if sdf_grid is not None else None
        self.layers = ([t.detach().requires_grad_(True) for t in layers]
                       if layers is not None else None)
        params = [self.data] + ([self.sdf] if self.sdf is not None else self.layers)
        self.opt = DeviceAdam(params, lr=lr)
        self.cam_buf = cam_buf
        self.n_cams, self.W, self.H = int(n_cams), int(width), int(height)
        self.interval, self.s = float(interval), float(s)
        self.mode = int(scatter_mode)
        self.gen = torch.Generator(device=data_grid.device).manual_seed(seed)

    def grid(self):
        if self.sdf is not None:
            return self.sdf
        return LayerToGrid.apply(*self.layers)

    def draw(self, batch):
        """int32 ray codes cam * W * H + pixel, uniform over all cameras."""
        n_all = self.n_cams * self.W * self.H
        return torch.randint(0, n_all, (batch,), generator=self.gen, device=self.data.device,
                             dtype=torch.int64).to(torch.int32)

    def state_dict(self):
        """Parameters, optimiser state and the batch generator (host
        copies): ``torch.save`` it, and ``load_state_dict`` resumes the run
        where it stopped."""
        params = {"data": self.data.detach().cpu()}
        if self.sdf is not None:
            params["sdf"] = self.sdf.detach().cpu()
        else:
            params["layers"] = [t.detach().cpu() for t in self.layers]
        return {"params": params, "adam": self.opt.state_dict(),
                "gen": self.gen.get_state(), "s": self.s}

    @torch.no_grad()
    def load_state_dict(self, sd):
        p = sd["params"]
        self.data.copy_(p["data"])
        if self.sdf is not None:
            self.sdf.copy_(p["sdf"])
        else:
            for dst, src in zip(self.layers, p["layers"]):
                dst.copy_(src)
        self.opt.load_state_dict(sd["adam"])
        self.gen.set_state(sd["gen"])
        self.s = float(sd["s"])

    def step(self, rays, target):
        """Forward + MSE + adjoint + Adam on ``rays`` (int32 codes) against
        ``target`` [n, 3].  Returns the loss as a device scalar (no sync)."""
        self.opt.zero_grad()
        c, _ = fused.FusedRender.apply(self.grid(), self.data, self.cam_buf, self.n_cams, self.W,
                                       self.H, rays, self.interval, self.s, self.mode)
        loss = ((c - target) ** 2).mean()
        loss.backward()
        self.opt.step()
        return loss.detach()
