"""Camera ray lists -- the caller contract of datasets/utils.py:5-56.

``RayCut`` once per camera (``GenerateRayPoints(mask)``), then per step
``GetRayPoints`` builds the five sample lists with ``GenerateRayPoints``.
Differences from the reference: the cut results stay on the GPU (no
D2H/H2D round trip per step) and the per-step lists are allocated with
``torch.empty`` since the kernel overwrites every row.
"""
from __future__ import annotations

import torch

from . import qsdvr


class Camera:
    def __init__(self, args: qsdvr.CameraArgs, interval: float = 0.03, reso: int = 128):
        self.args = args
        self.args.pixelCount = args.width * args.height
        self.interval = interval
        self.reso = reso

    def GenerateRayPoints(self, mask=None):
        P = self.args.pixelCount
        ray = torch.zeros((P, 4), dtype=torch.int32, device="cuda")
        ori = torch.zeros((P, 3), dtype=torch.float32, device="cuda")
        dirs = torch.zeros((P, 3), dtype=torch.float32, device="cuda")
        qsdvr.RayCut(self.args, mask, ray, ori, dirs, self.interval)
        if mask is not None:
            m = mask.reshape(-1).to("cuda")
            ray, ori, dirs = ray[m].contiguous(), ori[m].contiguous(), dirs[m].contiguous()
        self.rayList, self.originList, self.dirList = ray, ori, dirs

    def GetRayPoints(self) -> dict:
        a = self.args
        dev = "cuda"
        d = {
            "rayList": self.rayList,
            "sdfPointList": torch.empty((a.sdfCount, 3), dtype=torch.float32, device=dev),
            "sdfIndexList": torch.empty((a.sdfCount,), dtype=torch.int32, device=dev),
            "renderPointList": torch.empty((a.renderCount, 3), dtype=torch.float32, device=dev),
            "renderIndexList": torch.empty((a.renderCount,), dtype=torch.int32, device=dev),
            "viewDirList": torch.empty((a.renderCount, 3), dtype=torch.float32, device=dev),
        }
        qsdvr.GenerateRayPoints(a, d["sdfPointList"], d["sdfIndexList"], d["renderPointList"],
                                d["renderIndexList"], d["viewDirList"], self.rayList,
                                self.originList, self.dirList, self.interval, self.reso)
        return d
