"""Shared test helpers: tolerance metric and workload setup."""
from __future__ import annotations

import numpy as np
import torch


def scale_aware_close(x, ref, rtol):
    """Per element |x - ref| <= rtol * (|ref| + max|ref|) (vector rows use
    their norm).  Returns (ok, worst ratio, normwise relative error)."""
    x = torch.as_tensor(x).double().cpu()
    ref = torch.as_tensor(ref).double().cpu()
    assert x.shape == ref.shape, (x.shape, ref.shape)
    if ref.numel() == 0:
        return True, 0.0, 0.0
    if ref.dim() >= 2:
        err = torch.linalg.norm((x - ref).reshape(ref.shape[0], -1), dim=1)
        mag = torch.linalg.norm(ref.reshape(ref.shape[0], -1), dim=1)
    else:
        err = (x - ref).abs()
        mag = ref.abs()
    bound = rtol * (mag + mag.max())
    ratio = (err / bound.clamp_min(1e-300)).max().item()
    normwise = (torch.linalg.norm(x - ref) / torch.linalg.norm(ref).clamp_min(1e-300)).item()
    ok = bool(torch.isfinite(x).all()) and ratio <= 1.0
    return ok, ratio, normwise


def assert_close(x, ref, rtol, what=""):
    ok, ratio, nw = scale_aware_close(x, ref, rtol)
    assert ok, f"{what}: worst error / bound = {ratio:.3g} (normwise rel {nw:.3g}, rtol {rtol})"
    return ratio, nw
