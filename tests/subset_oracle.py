"""The float64 oracle on a subset of the rays of a full-size workload.

The grids of the largest workloads do not fit the CPU oracle in float64
(the 384^3 appearance grid alone is 12 GB in float32), so the oracle runs on
the rows the subset's samples touch: the cells of its SDF and render
samples and the 8 vertex rows of every render sample, gathered from the
device grids (``oracle.render.compact_index``).  The GPU renders the whole
workload; the loss is taken over the subset rays only, so the GPU gradient
must equal the oracle's at the gathered rows and be exactly zero elsewhere.
"""
from __future__ import annotations

import numpy as np
import torch

from oracle import raygen
from oracle import render as orc
from tests.helpers import assert_close

FWD, BWD = 1e-5, 1e-4


def strided(n_total, n_sub):
    """``n_sub`` ray positions spread evenly over ``n_total`` (fixed)."""
    step = n_total / n_sub
    return (np.arange(n_sub) * step + step / 2).astype(np.int64)


def concat_samples(parts):
    """Join sample dicts of consecutive ray groups into one (offsets
    shifted)."""
    out = {k: [] for k in ("rayList", "sdfPointList", "sdfIndexList", "renderPointList",
                           "renderIndexList", "viewDirList")}
    sb = rb = 0
    for p in parts:
        ray = p["rayList"].view(np.uint32).astype(np.int64).copy()
        ray[:, 1] += sb
        ray[:, 3] += rb
        out["rayList"].append(ray)
        for k in out:
            if k != "rayList":
                out[k].append(p[k])
        sb += p["sdfPointList"].shape[0]
        rb += p["renderPointList"].shape[0]
    res = {k: np.concatenate(v) for k, v in out.items()}
    res["rayList"] = res["rayList"].astype(np.uint32).view(np.int32)
    return res


def camera_samples(cams, codes, W, H, interval, R):
    """Oracle sample lists of the rays ``codes`` (cam * W * H + pixel, any
    order).  Returns (samples, order, centres): the samples hold the rays in
    the order ``codes[order]`` (grouped by camera), ``centres`` [S_r, 3] is
    the camera centre of every render sample."""
    codes = np.asarray(codes, dtype=np.int64)
    order = np.argsort(codes, kind="stable")
    sc = codes[order]
    cam_of = sc // (W * H)
    parts, centres = [], []
    for c in np.unique(cam_of):
        pix = sc[cam_of == c] - c * W * H
        smp = raygen.ray_samples(cams[c].as_dict(), np.float32(interval), R, pixels=pix)
        parts.append(smp)
        n_r = smp["renderPointList"].shape[0]
        centres.append(np.repeat(cams[c].translation.astype(np.float64)[None], n_r, 0))
    return concat_samples(parts), order, np.concatenate(centres)


def oracle_subset(sdf_grid, data_grid, smp, R, s, centres, target, depth_w=0.01):
    """Colour, depth and the gradients of
    mean((colour - target)^2) + depth_w * mean(depth) on the rays of ``smp``,
    in float64 on the touched rows.  Returns a dict with the touched
    ``cells`` and ``verts`` (sorted row numbers) and the results."""
    cells = np.unique(np.concatenate([smp["sdfIndexList"], smp["renderIndexList"]]).astype(np.int64))
    verts = np.unique(orc.vertex_rows(smp["renderIndexList"], R).numpy())
    dev = sdf_grid.device
    gc = sdf_grid.detach().reshape(-1, 7)[torch.as_tensor(cells, device=dev)].cpu().double()
    dc = data_grid.detach()[torch.as_tensor(verts, device=dev)].cpu().double()
    gc.requires_grad_(True)
    dc.requires_grad_(True)
    col, dep = orc.render(gc, dc, smp, R, s, with_depth=True, cam_center=centres,
                          cell_rows=cells, vertex_rows=verts)
    loss = orc.mse_loss(col, torch.as_tensor(target).double()) + depth_w * dep.mean()
    loss.backward()
    return {"cells": cells, "verts": verts, "colour": col.detach(), "depth": dep.detach(),
            "grad_sdf": gc.grad, "grad_data": dc.grad}


def gpu_subset_loss(colour, depth, rows, target, depth_w=0.01):
    """The same loss on the GPU's full-workload outputs, over ``rows``."""
    idx = torch.as_tensor(rows, device=colour.device)
    return ((colour[idx] - target.to(colour.device)) ** 2).mean() + depth_w * depth[idx].mean()


def check_subset(ref, colour_rows, depth_rows, grad_sdf, grad_data, what):
    """Colour and depth at 1e-5, gradients at 1e-4 at the touched rows, and
    both gradient grids exactly zero at every other row."""
    assert_close(colour_rows, ref["colour"], FWD, f"{what} colour")
    assert_close(depth_rows, ref["depth"], FWD, f"{what} depth")
    dev = grad_sdf.device
    cells = torch.as_tensor(ref["cells"], device=dev)
    verts = torch.as_tensor(ref["verts"], device=dev)
    gq = grad_sdf.reshape(-1, 7)
    assert_close(gq[cells].cpu(), ref["grad_sdf"], BWD, f"{what} sdf-grid grad")
    assert_close(grad_data[verts].cpu(), ref["grad_data"], BWD, f"{what} appearance grad")
    gq = gq.clone()
    gq.index_fill_(0, cells, 0.0)
    assert float(gq.abs().max()) == 0.0, f"{what}: sdf-grid gradient outside the touched cells"
    grad_data.index_fill_(0, verts, 0.0)  # in place: the caller's copy is not reused
    assert float(grad_data.abs().max()) == 0.0, f"{what}: appearance gradient outside touched rows"


def listed_lists(rs, t, R):
    """Sample lists for the oracle renderer from explicit positions t
    [n, K] along the rays of the ray state ``rs`` [n, 8] (origin, count
    bits, direction), with the kernels' float32 operation order:
    q = ((o + d t) + 1) R/2, render samples at 0.5 (t_j + t_{j+1})."""
    r = rs.cpu().numpy().astype(np.float32)
    cnt = rs[:, 3].contiguous().view(torch.int32).cpu().numpy()
    t = t.cpu().numpy().astype(np.float32)
    n, K = t.shape
    live = cnt >= 2
    half = np.float32(R) / np.float32(2.0)

    def cells(tt, rows):
        q = [((r[rows, a][:, None] + r[rows, 4 + a][:, None] * tt) + np.float32(1.0)) * half
             for a in range(3)]
        c = [np.floor(x).astype(np.int64) for x in q]
        loc = [x * np.float32(2.0) - ((2 * ci).astype(np.float32) + np.float32(1.0))
               for x, ci in zip(q, c)]
        idx = c[0] + c[1] * R + c[2] * R * R
        return np.stack(loc, -1).reshape(-1, 3), idx.reshape(-1)

    rows = np.nonzero(live)[0]
    sp, si = cells(t[rows], rows)
    mid = (np.float32(0.5) * (t[rows][:, :-1] + t[rows][:, 1:])).astype(np.float32)
    rp, ri = cells(mid, rows)
    sc = np.where(live, K, 0)
    rc = np.where(live, K - 1, 0)
    ray = np.stack([sc, np.cumsum(sc) - sc, rc, np.cumsum(rc) - rc], 1).astype(np.int32)
    vd = np.repeat(r[rows, 4:7], K - 1, axis=0)
    return {"rayList": ray, "sdfPointList": sp, "sdfIndexList": si.astype(np.int32),
            "renderPointList": rp, "renderIndexList": ri.astype(np.int32), "viewDirList": vd}
