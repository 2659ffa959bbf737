"""Bit-exact float32 restatement of the reference ray generation.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, operation by operation and in the same association order:
  * ``TransCameraToWorld``          kernel/camera.cu:3-10
  * ``RayCutKernel``                kernel/camera.cu:12-58
  * host scan inside ``RayCut``     kernel/camera.cu:111-127
  * ``GenerateRayPointsKernel``     kernel/camera.cu:60-96 (+ host :132-146)
  * ``vec3f`` arithmetic            include/common.h:25-73

Every arithmetic step is a numpy float32 operation on float32 operands, so
each step is one IEEE-754 round-to-nearest operation with no fused
multiply-add -- the same arithmetic the reference performs when compiled
without contraction.  Python float scalars are always converted with
``np.float32`` first so no step is silently widened to float64.

Pinned by (tests/test_oracle_raygen.py): a scalar per-pixel loop written
independently with Python ``struct`` float32 rounding (pins the
vectorisation), axis-aligned cameras whose sample counts are known in closed
form, float64 geometry (counts within one step of the exact chord length),
and the invariants offsets == exclusive scan of counts, local points in
[-1, 1), 0 <= cell index < R^3.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
CUBE = F32(0.9999)  # kernel/camera.cu:46


def _f(x) -> np.float32:
    return np.float32(x)


def camera_to_world(cam: dict, u: np.ndarray, v: np.ndarray, z) -> tuple:
    """kernel/camera.cu:3-10.  (u, v, z) pixel/depth -> world point.

    x = ((u + cx) / fx) * z ; y = ((v + cy) / fy) * z ;
    P = ((r0*x + r1*y) + r2*z) + t  (left to right, per row).
    """
    rot = np.asarray(cam["rotation"], dtype=F32).reshape(9)
    t = np.asarray(cam["translation"], dtype=F32).reshape(3)
    z = _f(z)
    x = ((u + _f(cam["cx"])) / _f(cam["fx"])) * z
    y = ((v + _f(cam["cy"])) / _f(cam["fy"])) * z
    px = ((rot[0] * x + rot[1] * y) + rot[2] * z) + t[0]
    py = ((rot[3] * x + rot[4] * y) + rot[5] * z) + t[1]
    pz = ((rot[6] * x + rot[7] * y) + rot[8] * z) + t[2]
    return px, py, pz


def ray_cut_kernel(cam: dict, interval, pixels=None) -> dict:
    """kernel/camera.cu:12-58, vectorised over all W*H pixels (or over the
    pixel ids ``pixels``, in that order: each pixel's ray depends on its id
    alone).

    Returns dict with float32 ``dir`` [P,3], ``origin`` [P,3] (first inside
    step; zeros when none) and uint32 ``count`` [P] (steps strictly inside
    the +-0.9999 cube).
    """
    W, H = int(cam["width"]), int(cam["height"])
    interval = _f(interval)
    if pixels is None:
        pid = np.arange(W * H, dtype=np.int64)
    else:
        pid = np.asarray(pixels, dtype=np.int64).reshape(-1)
        assert pid.size == 0 or (pid.min() >= 0 and pid.max() < W * H)
    P = pid.shape[0]
    u = (pid % W).astype(F32) + F32(0.5)
    v = (pid // W).astype(F32) + F32(0.5)
    nx, ny, nz = camera_to_world(cam, u, v, cam["near"])
    fx_, fy_, fz_ = camera_to_world(cam, u, v, cam["far"])
    dx, dy, dz = fx_ - nx, fy_ - ny, fz_ - nz
    nrm = np.sqrt((dx * dx + dy * dy) + dz * dz)           # common.h:66-69
    dirx, diry, dirz = dx / nrm, dy / nrm, dz / nrm          # common.h:58-65
    with np.errstate(divide="ignore", invalid="ignore"):
        ex = (fx_ - nx) / dirx
        ey = (fy_ - ny) / diry
        ez = (fz_ - nz) / dirz
    end = np.where(dirx != 0, ex, np.where(diry != 0, ey, ez)).astype(F32)
    cand = (end / interval).astype(np.uint32)                 # camera.cu:40
    count = np.zeros(P, dtype=np.uint32)
    ox = np.zeros(P, dtype=F32)
    oy = np.zeros(P, dtype=F32)
    oz = np.zeros(P, dtype=F32)
    jmax = int(cand.max()) if P else 0
    for j in range(jmax):
        fj = F32(j)
        live = cand > j
        px = nx + (dirx * fj) * interval                       # camera.cu:45
        py = ny + (diry * fj) * interval
        pz = nz + (dirz * fj) * interval
        inside = live & (px < CUBE) & (px > -CUBE) & (py < CUBE) & (py > -CUBE) \
            & (pz < CUBE) & (pz > -CUBE)
        first = inside & (count == 0)
        ox = np.where(first, px, ox)
        oy = np.where(first, py, oy)
        oz = np.where(first, pz, oz)
        count += inside.astype(np.uint32)
    return {
        "dir": np.stack([dirx, diry, dirz], 1).astype(F32),
        "origin": np.stack([ox, oy, oz], 1).astype(F32),
        "count": count,
    }


def ray_cut(cam: dict, mask: np.ndarray | None, interval, pixels=None):
    """kernel/camera.cu:98-130: kernel + host scan over masked pixels.

    Returns (rayList int32 [P,4], origin f32 [P,3], dir f32 [P,3],
    sdfTotal, renderTotal).  Rows of unmasked pixels carry the kernel's raw
    sdfCount in column 0 and zeros elsewhere (the caller zero-fills rayList,
    datasets/utils.py:43-45, and the reference only rewrites masked rows).
    With ``pixels`` the rows are those pixels' rays, in that order, and the
    scan runs over them alone.
    """
    k = ray_cut_kernel(cam, interval, pixels)
    P = k["count"].shape[0]
    if mask is None:
        mask = np.ones(P, dtype=bool)
    mask = np.asarray(mask, dtype=bool).reshape(P)
    cnt = k["count"].astype(np.int64)
    sdf = np.where(cnt < 2, 0, cnt)                             # camera.cu:117-120
    sdf_m = np.where(mask, sdf, 0)
    ren_m = np.where(sdf_m > 1, sdf_m - 1, 0)                   # camera.cu:124
    sdf_off = np.cumsum(sdf_m) - sdf_m                          # exclusive scan
    ren_off = np.cumsum(ren_m) - ren_m
    ray = np.zeros((P, 4), dtype=np.int64)
    ray[:, 0] = np.where(mask, sdf, cnt)
    ray[:, 1] = np.where(mask, sdf_off, 0)
    ray[:, 2] = np.where(mask, ren_m, 0)
    ray[:, 3] = np.where(mask, ren_off, 0)
    return (ray.astype(np.uint32).view(np.int32), k["origin"], k["dir"],
            int(sdf_m.sum()), int(ren_m.sum()))


def _points(ori, dirs, fj_vec, interval, R, render: bool):
    vo = F32(R) / F32(2.0)                                      # camera.cu:146
    if render:
        s = (fj_vec + F32(0.5)) * interval                      # camera.cu:86
        q = [((ori[:, a] + dirs[:, a] * s) + F32(1.0)) * vo for a in range(3)]
    else:
        q = [((ori[:, a] + (dirs[:, a] * fj_vec) * interval) + F32(1.0)) * vo
             for a in range(3)]                                 # camera.cu:73
    cell = [np.floor(c).astype(np.int64) for c in q]            # camera.cu:74-76
    local = [(q[a] * F32(2.0)) - ((2 * cell[a]).astype(F32) + F32(1.0))
             for a in range(3)]                                 # camera.cu:77
    index = cell[0] + cell[1] * R + cell[2] * (R * R)           # camera.cu:78
    return np.stack(local, 1).astype(F32), index


def generate_ray_points(rayList, origin, dirs, interval, R):
    """kernel/camera.cu:60-96 over the rows of ``rayList`` (masked rays).

    Returns sdfPoint f32[S_sdf,3], sdfIndex i32[S_sdf], renderPoint
    f32[S_r,3], renderIndex i32[S_r], viewDir f32[S_r,3].
    """
    interval = _f(interval)
    ray = np.asarray(rayList).view(np.uint32).astype(np.int64)
    ori = np.asarray(origin, dtype=F32)
    dirs = np.asarray(dirs, dtype=F32)
    out = {}
    for kind, ccol, ocol, render in (("sdf", 0, 1, False), ("render", 2, 3, True)):
        cnt = ray[:, ccol]
        total = int((ray[:, ocol] + cnt).max()) if len(cnt) and cnt.max() > 0 else 0
        rid = np.repeat(np.arange(len(cnt)), cnt)
        j = np.arange(int(cnt.sum()), dtype=np.int64) - np.repeat(np.cumsum(cnt) - cnt, cnt)
        dst = ray[rid, ocol] + j
        pts, idx = _points(ori[rid], dirs[rid], j.astype(F32), interval, R, render)
        P_ = np.zeros((total, 3), dtype=F32)
        I_ = np.zeros(total, dtype=np.int64)
        P_[dst] = pts
        I_[dst] = idx
        out[kind + "Point"] = P_
        out[kind + "Index"] = I_.astype(np.int32)
        if render:
            V_ = np.zeros((total, 3), dtype=F32)
            V_[dst] = dirs[rid]
            out["viewDir"] = V_
    return (out["sdfPoint"], out["sdfIndex"], out["renderPoint"],
            out["renderIndex"], out["viewDir"])


def ray_samples(cam: dict, interval, R, mask=None, pixels=None) -> dict:
    """RayCut followed by GenerateRayPoints on the masked rows, as the
    reference's Camera does (datasets/utils.py:16-56); with ``pixels``, on
    those pixels' rays only (in that order)."""
    ray, ori, dirs, st, rt = ray_cut(cam, mask, interval, pixels)
    P = ray.shape[0]
    m = np.ones(P, dtype=bool) if mask is None else np.asarray(mask, bool).reshape(P)
    ray_m, ori_m, dir_m = ray[m], ori[m], dirs[m]
    sp, si, rp, ri, vd = generate_ray_points(ray_m, ori_m, dir_m, interval, R)
    assert sp.shape[0] == st and rp.shape[0] == rt
    return {"rayList": ray_m, "origin": ori_m, "dir": dir_m,
            "sdfPointList": sp, "sdfIndexList": si, "renderPointList": rp,
            "renderIndexList": ri, "viewDirList": vd,
            "sdfCount": st, "renderCount": rt}
