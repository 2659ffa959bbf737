"""Autograd layer over the op surface -- the caller contract of
models/render.py:8-237, rebuilt on the HIP library.

Class names, ``forward`` arguments, saved tensors and output/gradient
conventions follow the reference so code written against ``models.render``
runs unchanged; every op goes through ``qsdvr_amd.qsdvr`` (the C-ABI).
"""
from __future__ import annotations

import math
from typing import Any

import torch
from torch import autograd, nn

from . import qsdvr


class GridInterpolation(autograd.Function):
    """models/render.py:8-35.  No gradient to the points (render.py:35)."""

    @staticmethod
    def forward(ctx: Any, dataGrid, pointList, indexList):
        n = pointList.size(0)
        Cc = dataGrid.size(-1)
        reso = dataGrid.size(0) if dataGrid.dim() == 4 else round(
            (dataGrid.numel() // Cc) ** (1.0 / 3.0))
        out = torch.empty((n, Cc), dtype=torch.float32, device=pointList.device)
        qsdvr.GridInterpolationForward(dataGrid, pointList, indexList, out, reso)
        ctx.save_for_backward(pointList, indexList)
        ctx.reso = reso
        ctx.shape = dataGrid.shape
        return out

    @staticmethod
    def backward(ctx: Any, gradOutput):
        pointList, indexList = ctx.saved_tensors
        g = torch.zeros(ctx.shape, dtype=torch.float32, device=pointList.device)
        qsdvr.GridInterpolationBackward(g, pointList, indexList, gradOutput.contiguous(),
                                        ctx.reso)
        return g, None, None


class Shader(autograd.Function):
    """models/render.py:38-65.  No gradient to the view directions."""

    @staticmethod
    def forward(ctx: Any, normalList, viewDirList, dataList):
        n = normalList.size(0)
        out = torch.empty((n, 3), dtype=torch.float32, device=normalList.device)
        spec = torch.empty((n, 3), dtype=torch.float32, device=normalList.device)
        qsdvr.ShaderForward(out, normalList, viewDirList, dataList, spec)
        ctx.save_for_backward(normalList, viewDirList, dataList, spec)
        return out

    @staticmethod
    def backward(ctx: Any, gradOutput):
        normalList, viewDirList, dataList, spec = ctx.saved_tensors
        n = normalList.size(0)
        gN = torch.empty((n, 3), dtype=torch.float32, device=normalList.device)
        gD = torch.empty_like(dataList)
        qsdvr.ShaderBackward(gradOutput.contiguous(), normalList, viewDirList, dataList, spec,
                             gN, gD)
        return gN, None, gD


class RayAggregate(autograd.Function):
    """models/render.py:68-98."""

    @staticmethod
    def forward(ctx: Any, rgbList, sdfList, rayList, logisticCoef):
        rays = rayList.size(0)
        nr = rgbList.size(0)
        out = torch.empty((rays, 3), dtype=torch.float32, device=rgbList.device)
        alpha = torch.empty(nr, dtype=torch.float32, device=rgbList.device)
        T = torch.empty(nr, dtype=torch.float32, device=rgbList.device)
        qsdvr.RayAggregateForward(out, rgbList, sdfList, rayList, alpha, T, logisticCoef)
        ctx.save_for_backward(rgbList, sdfList, rayList, alpha, T)
        ctx.logisticCoef = logisticCoef
        return out

    @staticmethod
    def backward(ctx: Any, gradOutput):
        rgbList, sdfList, rayList, alpha, T = ctx.saved_tensors
        gR = torch.empty((rgbList.size(0), 3), dtype=torch.float32, device=rgbList.device)
        gS = torch.empty(sdfList.size(0), dtype=torch.float32, device=rgbList.device)
        qsdvr.RayAggregateBackward(gradOutput.contiguous(), rgbList, sdfList, rayList, alpha, T,
                                   gR, gS, ctx.logisticCoef)
        return gR, gS, None, None


class LayerToGrid(autograd.Function):
    """models/render.py:101-122.  Layer gradients come back on the layers'
    device (CPU for the reference's CPU parameters)."""

    @staticmethod
    def forward(ctx: Any, xLayer, yLayer, zLayer, offset):
        reso = xLayer.size(0)
        out = torch.empty((reso, reso, reso, 7), dtype=torch.float32, device="cuda")
        qsdvr.LayerToGridForward(out, xLayer.detach().contiguous(), yLayer.detach().contiguous(),
                                 zLayer.detach().contiguous(), offset.detach().contiguous())
        ctx.reso = reso
        ctx.dev = xLayer.device
        return out

    @staticmethod
    def backward(ctx: Any, gradOutput):
        R, dev = ctx.reso, ctx.dev
        gx = torch.empty(R, dtype=torch.float32, device=dev)
        gy = torch.empty(R, dtype=torch.float32, device=dev)
        gz = torch.empty(R, dtype=torch.float32, device=dev)
        go = torch.empty(4, dtype=torch.float32, device=dev)
        qsdvr.LayerToGridBackward(gradOutput.contiguous(), gx, gy, gz, go)
        return gx, gy, gz, go


class SampleSDF(autograd.Function):
    """models/render.py:125-146."""

    @staticmethod
    def forward(ctx: Any, sdfGrid, pointList, indexList):
        out = torch.empty(pointList.size(0), dtype=torch.float32, device=pointList.device)
        qsdvr.SampleSDFForward(out, sdfGrid, pointList, indexList, sdfGrid.size(0))
        ctx.save_for_backward(sdfGrid, pointList, indexList)
        ctx.reso = sdfGrid.size(0)
        return out

    @staticmethod
    def backward(ctx: Any, gradOutput):
        sdfGrid, pointList, indexList = ctx.saved_tensors
        g = torch.zeros_like(sdfGrid)
        qsdvr.SampleSDFBackward(gradOutput.contiguous(), sdfGrid, pointList, indexList, g,
                                ctx.reso)
        return g, None, None


class SampleNormal(autograd.Function):
    """models/render.py:149-169."""

    @staticmethod
    def forward(ctx: Any, sdfGrid, pointList, indexList):
        out = torch.empty((pointList.size(0), 3), dtype=torch.float32, device=pointList.device)
        qsdvr.SampleNormalForward(out, sdfGrid, pointList, indexList)
        ctx.save_for_backward(sdfGrid, pointList, indexList)
        return out

    @staticmethod
    def backward(ctx: Any, gradOutput):
        sdfGrid, pointList, indexList = ctx.saved_tensors
        g = torch.zeros_like(sdfGrid)
        qsdvr.SampleNormalBackward(gradOutput.contiguous(), sdfGrid, pointList, indexList, g)
        return g, None, None


class QuadricGrid(nn.Module):
    """models/render.py:172-206: separable layer parameterisation (CPU
    parameters, as in the reference) turned into the quadric grid on the GPU.
    ``init_as_sphere`` follows the oracle's sphere (render_torch.py:22-28),
    the one that actually describes a sphere of the given radius."""

    def __init__(self, reso: int = 128, device="cpu"):
        super().__init__()
        self.reso = reso
        self.xLayer = nn.Parameter(torch.ones(reso, dtype=torch.float32, device=device))
        self.yLayer = nn.Parameter(torch.ones(reso, dtype=torch.float32, device=device))
        self.zLayer = nn.Parameter(torch.ones(reso, dtype=torch.float32, device=device))
        off = torch.zeros(4, dtype=torch.float32, device=device)
        self.offset = nn.Parameter(off)
        self.init_as_sphere()

    @torch.no_grad()
    def init_as_sphere(self, radius=1.0):
        r = radius * self.reso
        c = 1 - self.reso
        self.offset[0] = 2 * c
        self.offset[1] = 2 * c
        self.offset[2] = 2 * c
        self.offset[3] = 3 * c * c - r * r

    def grid(self):
        return LayerToGrid.apply(self.xLayer, self.yLayer, self.zLayer, self.offset)

    def forward(self, renderPointList, renderIndexList, sdfPointList, sdfIndexList):
        sdfGrid = self.grid()
        sdfList = SampleSDF.apply(sdfGrid, sdfPointList, sdfIndexList)
        normalList = SampleNormal.apply(sdfGrid, renderPointList, renderIndexList)
        return sdfList, normalList


def render_lists(sdfGrid, renderData, input: dict, logisticCoef: float):
    """The composition of models/render.py:225-234 on given grids:
    SampleSDF -> SampleNormal -> GridInterpolation -> Shader -> RayAggregate."""
    sdfList = SampleSDF.apply(sdfGrid, input["sdfPointList"], input["sdfIndexList"])
    normalList = SampleNormal.apply(sdfGrid, input["renderPointList"], input["renderIndexList"])
    datalist = GridInterpolation.apply(renderData, input["renderPointList"],
                                       input["renderIndexList"])
    rgbList = Shader.apply(normalList, input["viewDirList"], datalist)
    return RayAggregate.apply(rgbList, sdfList, input["rayList"], logisticCoef)


class RenderGrid(nn.Module):
    """models/render.py:209-237."""

    def __init__(self, reso: int = 128, logisticCoef: float = 5, channels: int = 32):
        super().__init__()
        self.quadricGrid = QuadricGrid(reso)
        self.reso = reso
        data = torch.zeros((reso + 1, reso + 1, reso + 1, channels), dtype=torch.float32,
                           device="cuda")
        nc = (channels - 5) // 3
        data[..., 3 * nc + 4] = 1
        data[..., 3 * nc + 3] = 0.1
        data[..., 3 * nc:3 * nc + 3] = 1
        data[..., 0:3] = 0.5
        self.renderData = nn.Parameter(data)
        self.logisticCoef = logisticCoef

    def forward(self, input):
        sdfList, normalList = self.quadricGrid.forward(
            input["renderPointList"], input["renderIndexList"], input["sdfPointList"],
            input["sdfIndexList"])
        datalist = GridInterpolation.apply(self.renderData, input["renderPointList"],
                                           input["renderIndexList"])
        rgbList = Shader.apply(normalList, input["viewDirList"], datalist)
        return RayAggregate.apply(rgbList, sdfList, input["rayList"], self.logisticCoef)

    def set_logisticCoef(self, logisticCoef):
        self.logisticCoef = logisticCoef
