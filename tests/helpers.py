"""Shared test helpers: the tolerance metric.

A result x is close to the float64 reference ref at relative tolerance rtol
when both hold:

* per element (vector rows by their norm):
  |x - ref| <= rtol * (|ref| + floor * max|ref|)
  -- relative to the element itself, with an absolute floor a small fraction
  of the tensor's largest entry (gradient entries that cancel to near zero
  carry float32 rounding of their larger terms, which no purely relative
  test can admit);
* norm-wise: ||x - ref|| <= rtol * ||ref||.

The floor depends on what is compared (``floor_for``):
* values (colours, depths, per-sample outputs; rtol 1e-5): 0.25.  Float32
  compositing cannot do better per element on rays whose colour cancels
  (SH shading is signed): the reference's own formulas evaluated in float32
  (the oracle run in float32) already need a floor of 0.03-0.05 on full
  128^3 frames, from alpha = (e_{i+1} - e_i) / (1 + e_{i+1}) subtracting
  nearly equal exponentials and T carried as a product over ~100 samples;
  the GPU kernels (with 1-ulp hardware rsq/rcp in the SDF and compositing)
  measured at most 0.18 on the same rays.
* gradients (rtol 1e-4): 0.01.
* bfloat16 proposal outputs (rtol 3e-3): 1.0 (``BF16_ATOL_FRAC``): an
  activation within float32 rounding of a bf16 rounding boundary may round
  the other way, which moves a logit by one bf16 ulp of that activation
  times a weight, however small the logit.
"""
from __future__ import annotations

import json
import os

import torch

# absolute floor of the per-element bound, as a fraction of max|ref|
VALUE_ATOL_FRAC = 0.25
GRAD_ATOL_FRAC = 0.01
BF16_ATOL_FRAC = 1.0


def floor_for(rtol):
    """The per-element floor of a check at ``rtol``: values are compared at
    1e-5, gradients at 1e-4 and bf16 outputs at 3e-3 throughout the tests."""
    if rtol <= 1e-5:
        return VALUE_ATOL_FRAC
    if rtol <= 1e-4:
        return GRAD_ATOL_FRAC
    return BF16_ATOL_FRAC


def scale_aware_close(x, ref, rtol, atol_frac=None):
    """Returns (ok, worst per-element ratio, normwise relative error,
    smallest atol_frac that would pass the per-element test)."""
    if atol_frac is None:
        atol_frac = floor_for(rtol)
    x = torch.as_tensor(x).double().cpu()
    ref = torch.as_tensor(ref).double().cpu()
    assert x.shape == ref.shape, (x.shape, ref.shape)
    if ref.numel() == 0:
        return True, 0.0, 0.0, 0.0
    if ref.dim() >= 2:
        err = torch.linalg.norm((x - ref).reshape(ref.shape[0], -1), dim=1)
        mag = torch.linalg.norm(ref.reshape(ref.shape[0], -1), dim=1)
    else:
        err = (x - ref).abs()
        mag = ref.abs()
    top = mag.max().clamp_min(1e-300)
    bound = rtol * (mag + atol_frac * top)
    ratio = (err / bound.clamp_min(1e-300)).max().item()
    need = ((err / rtol - mag) / top).max().clamp_min(0.0).item()
    normwise = (torch.linalg.norm(x - ref) / torch.linalg.norm(ref).clamp_min(1e-300)).item()
    finite = bool(torch.isfinite(x).all())
    ok = finite and ratio <= 1.0 and normwise <= rtol
    return ok, ratio, normwise, need


def assert_close(x, ref, rtol, what="", atol_frac=None):
    if atol_frac is None:
        atol_frac = floor_for(rtol)
    ok, ratio, nw, need = scale_aware_close(x, ref, rtol, atol_frac)
    log = os.environ.get("QSDVR_TOL_LOG")
    if log:
        with open(log, "a") as f:
            f.write(json.dumps({"what": what, "rtol": rtol, "ratio": ratio, "normwise": nw,
                                "atol_frac_needed": need,
                                "test": os.environ.get("PYTEST_CURRENT_TEST", "")}) + "\n")
    assert ok, (f"{what}: worst error / bound = {ratio:.3g}, normwise rel {nw:.3g} "
                f"(rtol {rtol}, floor {atol_frac} max|ref|; floor needed {need:.3g})")
    return ratio, nw
