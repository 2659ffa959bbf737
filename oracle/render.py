"""Restatement of the reference's pure-torch renderer (models/render_torch.py).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Device-agnostic torch-CPU code, generic in dtype: float32 reproduces the
reference's own arithmetic, float64 is the high-precision oracle the GPU
results are judged against.  Gradients are torch autograd of these forward
functions ("render_torch's autograd gradient").

Semantic contract (follow render_torch.py; fix only what stops it running):
  * interpolation weights are the raw local point p in [-1, 1) and the
    neighbour rows are i, i+1, i+R, i+R+1, i+R^2, ... (render_torch.py:128-143)
  * SH coefficient k, colour c lives in channel 3k+c (render_torch.py:147)
  * normal = F.normalize (eps 1e-12)                 (render_torch.py:100)
  * alpha = clamp(., 0, 10) with torch's clamp gradient (passes at 0)
  * every ray's colour goes to its own row (the reference writes image[0],
    render_torch.py:167 -- run-blocking defect fixed); empty rays are black.

Pinned by (tests/test_oracle_render.py): the closed-form sphere grid from
``init_as_sphere`` (render_torch.py:22-28), plane special cases for SDF and
normal, constant / indicator grids for interpolation and the transpose
(dot-product) test for its adjoint, real-SH closed forms and the
metallic=0 / coef0-only shading cases, for degree 3 (53 channels, no
reference counterpart) orthonormality on the sphere, harmonicity and the
textbook z-pole table under the axis substitution that maps it onto the
reference's degree-2 table, single-step SDF compositing and the
invariants T non-increasing / sum of weights <= 1, a literal per-ray Python
loop for compositing, and float64 central finite differences for every
gradient.
"""
from __future__ import annotations

import math

import numpy as np
import torch

# render_torch.py:181-185 -- created as a float32 tensor, so these are the
# float32 roundings of the decimal literals.
SH_FACTOR_F32 = np.array([0.2820947917, 0.4886025119, 0.4886025119, 0.4886025119,
                          1.0925484305, 1.0925484305, 0.3153915652, 1.0925484305,
                          0.5462742152], dtype=np.float32)

# Degree 3 (no reference counterpart): the unsigned real SH of degree 3 in
# the same y-pole family and m = -3..3 order as the reference's degree <= 2
# table, normalisations sqrt(35/32pi), sqrt(105/4pi), sqrt(21/32pi),
# sqrt(7/16pi), sqrt(21/32pi), sqrt(105/16pi), sqrt(35/32pi) as float32 of
# their 10-digit decimals.
SH3_FACTOR_F32 = np.array([0.5900435899, 2.8906114426, 0.4570457995, 0.3731763326,
                           0.4570457995, 1.4453057213, 0.5900435899], dtype=np.float32)
SH_FACTOR16_F32 = np.concatenate([SH_FACTOR_F32, SH3_FACTOR_F32])


def layers_to_grid(x, y, z, off):
    """render_torch.py:30-75 (== kernel/surface.cu:148-229 recurrences).

    Returns the quadric grid [R, R, R, 7] indexed [z][y][x][k] with
    A=X[x], B=Y[y], C=Z[z], D=X3[x], E=Y4[y], F=Z5[z],
    G = off[3] + X6[x] + Y6[y] + Z6[z].
    Differentiable in all inputs (autograd gives LayerToGrid's adjoint).
    """
    R = x.shape[0]

    def rec(a, o):
        d = [o]
        for i in range(1, R):
            d.append(2 * a[i - 1] + 2 * a[i] + d[i - 1])
        c = [torch.zeros((), dtype=a.dtype)]
        for i in range(1, R):
            c.append(3 * a[i - 1] + a[i] + 2 * d[i - 1] + c[i - 1])
        return torch.stack(d), torch.stack(c)

    x3, x6 = rec(x, off[0])
    y4, y6 = rec(y, off[1])
    z5, z6 = rec(z, off[2])
    one = torch.ones((R, R, R), dtype=x.dtype)
    X = lambda v: one * v.view(1, 1, R)     # noqa: E731 -- varies along x (last)
    Y = lambda v: one * v.view(1, R, 1)     # noqa: E731
    Z = lambda v: one * v.view(R, 1, 1)     # noqa: E731
    g = off[3] + X(x6) + Y(y6) + Z(z6)
    return torch.stack([X(x), Y(y), Z(z), X(x3), Y(y4), Z(z5), g], dim=-1)


def sphere_layers(R: int, radius: float = 1.0, dtype=torch.float64):
    """``QuadricGridTorch.init_as_sphere`` (render_torch.py:8-28): layers all
    ones and offset = (2c, 2c, 2c, 3c^2 - (radius*R)^2) with c = 1 - R."""
    rr = radius * R
    c = 1 - R
    ones = torch.ones(R, dtype=dtype)
    off = torch.tensor([2 * c, 2 * c, 2 * c, 3 * c * c - rr * rr], dtype=dtype)
    return ones.clone(), ones.clone(), ones.clone(), off


def compact_index(rows, idx):
    """Position of each full-grid row number of ``idx`` in ``rows``, the
    sorted row numbers a compacted grid holds (raises if one is absent).

    A compacted grid is the rows ``rows`` of a full grid gathered in that
    order; it lets the oracle run on the rows a few rays touch of a grid too
    large to convert to float64 whole.  The arithmetic is unchanged: only the
    row lookup goes through this map."""
    rows = torch.as_tensor(rows, dtype=torch.int64)
    flat = idx.reshape(-1).long()
    pos = torch.searchsorted(rows, flat)
    ok = (pos < rows.numel()) & (rows[pos.clamp(max=max(rows.numel() - 1, 0))] == flat)
    if not bool(ok.all()):
        raise KeyError("row missing from the compacted grid")
    return pos.reshape(idx.shape)


def _gather7(grid, idx, rows=None):
    g = grid.reshape(-1, 7)
    i = idx.long() if rows is None else compact_index(rows, idx)
    return g[i]


def sample_sdf(grid, pts, idx, R, rows=None):
    """render_torch.py:77-90 (kernel/surface.cu:54-72): f/||grad f||/R.
    ``rows``: the cell numbers held by a compacted ``grid`` (else the full
    grid [R,R,R,7])."""
    q = _gather7(grid, idx, rows)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    a0, a1, a2 = q[:, 0] * x, q[:, 1] * y, q[:, 2] * z
    a3, a4, a5, a6 = q[:, 3], q[:, 4], q[:, 5], q[:, 6]
    f = (a0 + a3) * x + (a1 + a4) * y + (a2 + a5) * z + a6
    n = torch.sqrt((2 * a0 + a3) * (2 * a0 + a3) + (2 * a1 + a4) * (2 * a1 + a4)
                   + (2 * a2 + a5) * (2 * a2 + a5))
    return f / n / R


def sample_normal(grid, pts, idx, rows=None):
    """render_torch.py:91-100: F.normalize(grad f) with eps 1e-12."""
    q = _gather7(grid, idx, rows)
    g = torch.stack([2 * q[:, 0] * pts[:, 0] + q[:, 3],
                     2 * q[:, 1] * pts[:, 1] + q[:, 4],
                     2 * q[:, 2] * pts[:, 2] + q[:, 5]], dim=1)
    return torch.nn.functional.normalize(g, dim=1)


def _lerp(a, b, w):
    """render_torch.py:170-172: (1 - w) * a + b * w."""
    w = w.reshape(-1, 1)
    return (1 - w) * a + b * w


def vertex_rows(idx, R):
    """The 8 appearance rows [S, 8] a sample of cell ``idx`` reads:
    i, i+1, i+R, i+R+1, i+R^2, i+R^2+1, i+R^2+R, i+R^2+R+1."""
    i = torch.as_tensor(idx).long().reshape(-1, 1)
    off = torch.tensor([0, 1, R, R + 1, R * R, R * R + 1, R * R + R, R * R + R + 1])
    return i + off


def grid_interpolation(data, pts, idx, R, rows=None):
    """render_torch.py:128-143.  data [(R+1)^3, C]; rows addressed with the
    cell's flat index i = x + yR + zR^2 and strides 1, R, R^2; weights are the
    raw local coordinates.  ``rows``: the vertex numbers held by a compacted
    ``data``."""
    d = data.reshape(-1, data.shape[-1])
    i000 = idx.long()
    i010 = i000 + R
    i100 = i000 + R * R
    i110 = i100 + R
    if rows is None:
        row = lambda i: d[i]                                      # noqa: E731
    else:
        row = lambda i: d[compact_index(rows, i)]                 # noqa: E731
    a00 = _lerp(row(i000), row(i000 + 1), pts[:, 0])
    a01 = _lerp(row(i010), row(i010 + 1), pts[:, 0])
    a0 = _lerp(a00, a01, pts[:, 1])
    a10 = _lerp(row(i100), row(i100 + 1), pts[:, 0])
    a11 = _lerp(row(i110), row(i110 + 1), pts[:, 0])
    a1 = _lerp(a10, a11, pts[:, 1])
    return _lerp(a0, a1, pts[:, 2])


def sh_basis(v, nc=9):
    """The basis polynomials without constants.  nc = 9: the degree-2 terms
    of render_torch.py:186-195, [1, x, y, z, xz, xy, 2y^2 - z^2 - x^2, yz,
    z^2 - x^2]; nc = 16 appends degree 3, [x(3z^2 - x^2), xyz,
    x(4y^2 - x^2 - z^2), y(2y^2 - 3x^2 - 3z^2), z(4y^2 - x^2 - z^2),
    y(z^2 - x^2), z(z^2 - 3x^2)]."""
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    zz, xx = z * z, x * x
    terms = [torch.ones_like(x), x, y, z, x * z, x * y, 2.0 * y * y - zz - xx, y * z, zz - xx]
    if nc == 16:
        yy = y * y
        terms += [x * (3.0 * zz - xx), x * y * z, x * (4.0 * yy - xx - zz),
                  y * (2.0 * yy - 3.0 * xx - 3.0 * zz), z * (4.0 * yy - xx - zz), y * (zz - xx),
                  z * (zz - 3.0 * xx)]
    elif nc != 9:
        raise ValueError("nc must be 9 or 16")
    return torch.stack(terms, dim=1)


def sh_irradiance(v, coef, factors=None):
    """render_torch.py:174-195.  v [S,3], coef [S,NC,3] -> [S,3] (NC = 9, or
    16 for degree 3)."""
    nc = coef.shape[1]
    default = SH_FACTOR_F32 if nc == 9 else SH_FACTOR16_F32
    K = torch.as_tensor(default if factors is None else factors, dtype=v.dtype)
    Y = sh_basis(v, nc) * K
    return (coef * Y.unsqueeze(-1)).sum(dim=1)


def shade(normal, view, data):
    """render_torch.py:144-152.  Returns (rgb [S,3], specularL [S,3]).
    Channels (C = 3 NC + 5): [3k+c] SH coef k colour c (k < NC), then
    diffuse RGB, metallic, ao; C = 32 (NC = 9) as in the reference, or 53
    (NC = 16, degree 3)."""
    nc = (data.shape[1] - 5) // 3
    ao = data[:, 3 * nc + 4:3 * nc + 5]
    diffuse = data[:, 3 * nc:3 * nc + 3]
    metallic = data[:, 3 * nc + 3:3 * nc + 4]
    specular = data[:, :3 * nc].reshape(-1, nc, 3)
    vdotn = (view * normal).sum(dim=1).reshape(-1, 1)
    reflect = view - normal * (2.0 * vdotn)
    L = sh_irradiance(reflect, specular)
    color = diffuse * (1 - metallic) + L * metallic
    return color * ao, L


def _ray_tables(rayList):
    ray = torch.as_tensor(np.asarray(rayList).view(np.uint32).astype(np.int64))
    return ray[:, 0], ray[:, 1], ray[:, 2], ray[:, 3]


def composite(rgb, sdf, rayList, s, t=None, return_weights=False):
    """render_torch.py:155-167, vectorised over rays (padding to the longest
    ray).  alpha_i = clamp((e_{i+1} - e_i) / (1 + e_{i+1}), 0, 10) with
    e = exp(-s * sdf); C = sum T_i alpha_i rgb_i, T_{i+1} = T_i (1 - alpha_i).

    ``t`` (optional [S_r]) adds the depth D = sum T_i alpha_i t_i.
    Returns colour [P,3] (and depth [P], weights [P,L] if asked)."""
    sc, so, rc, ro = _ray_tables(rayList)
    P = sc.shape[0]
    L = int(rc.max()) if P else 0
    dt = rgb.dtype
    if L == 0:
        col = torch.zeros((P, 3), dtype=dt)
        out = [col]
        if t is not None:
            out.append(torch.zeros(P, dtype=dt))
        if return_weights:
            out.append(torch.zeros((P, 0), dtype=dt))
        return out[0] if len(out) == 1 else tuple(out)
    j = torch.arange(L)
    valid = j.unsqueeze(0) < rc.unsqueeze(1)
    si0 = torch.where(valid, so.unsqueeze(1) + j, torch.zeros_like(valid, dtype=torch.long))
    si1 = torch.where(valid, so.unsqueeze(1) + j + 1, torch.zeros_like(si0))
    ri = torch.where(valid, ro.unsqueeze(1) + j, torch.zeros_like(si0))
    e0 = torch.exp(-s * sdf[si0])
    e1 = torch.exp(-s * sdf[si1])
    alpha = torch.clamp((e1 - e0) / (1.0 + e1), 0.0, 10.0)
    alpha = torch.where(valid, alpha, torch.zeros_like(alpha))
    one_m = 1 - alpha
    T = torch.cat([torch.ones((P, 1), dtype=dt), torch.cumprod(one_m, dim=1)[:, :-1]], dim=1)
    w = T * alpha
    col = (w.unsqueeze(-1) * rgb[ri]).sum(dim=1)
    out = [col]
    if t is not None:
        out.append((w * t[ri]).sum(dim=1))
    if return_weights:
        out.append(w)
    return out[0] if len(out) == 1 else tuple(out)


def composite_literal(rgb, sdf, rayList, s):
    """render_torch.py:155-167 as written (per-ray, per-sample Python loop),
    with the image[0] defect fixed and empty rays left black.  Small inputs
    only."""
    sc, so, rc, ro = _ray_tables(rayList)
    P = sc.shape[0]
    rows = []
    for r in range(P):
        color = torch.zeros(3, dtype=rgb.dtype)
        if int(rc[r]) > 0:
            e_i = torch.exp(-s * sdf[int(so[r])])
            T = torch.ones((), dtype=rgb.dtype)
            for i in range(int(rc[r])):
                e_i1 = torch.exp(-s * sdf[int(so[r]) + i + 1])
                alpha = torch.clamp((e_i1 - e_i) / (1.0 + e_i1), 0.0, 10.0)
                color = color + rgb[int(ro[r]) + i] * alpha * T
                e_i = e_i1
                T = T * (1 - alpha)
        rows.append(color)
    return torch.stack(rows) if rows else torch.zeros((0, 3), dtype=rgb.dtype)


def render(sdf_grid, data, samples: dict, R: int, s: float, with_depth=False, cam_center=None,
           cell_rows=None, vertex_rows=None):
    """RenderGridTorch.forward (render_torch.py:117-168) on given grids.

    ``samples`` holds the ray lists (rayList, sdfPointList, sdfIndexList,
    renderPointList, renderIndexList, viewDirList) as numpy arrays; grids are
    torch tensors of the caller's dtype (gradients flow to them).  With
    ``cell_rows`` / ``vertex_rows`` the grids are compacted (``compact_index``).
    ``cam_center`` is one camera centre [3] or one per render sample [S_r, 3]."""
    dt = sdf_grid.dtype
    sp = torch.as_tensor(samples["sdfPointList"]).to(dt)
    si = torch.as_tensor(samples["sdfIndexList"]).long()
    rp = torch.as_tensor(samples["renderPointList"]).to(dt)
    ri = torch.as_tensor(samples["renderIndexList"]).long()
    vd = torch.as_tensor(samples["viewDirList"]).to(dt)
    sdf = sample_sdf(sdf_grid, sp, si, R, cell_rows)
    nrm = sample_normal(sdf_grid, rp, ri, cell_rows)
    dl = grid_interpolation(data, rp, ri, R, vertex_rows)
    rgb, _ = shade(nrm, vd, dl)
    if not with_depth:
        return composite(rgb, sdf, samples["rayList"], s)
    t = render_sample_depth(samples, R, cam_center, dt)
    return composite(rgb, sdf, samples["rayList"], s, t=t)


def render_sample_depth(samples: dict, R: int, cam_center, dtype=torch.float64):
    """Distance from the camera centre to every render sample, with the
    sample's cube position rebuilt from its cell index and local point:
    x = (2*cell + 1 + p) / R - 1."""
    ri = torch.as_tensor(samples["renderIndexList"]).long()
    rp = torch.as_tensor(samples["renderPointList"]).to(dtype)
    cx = ri % R
    cy = (ri // R) % R
    cz = ri // (R * R)
    cell = torch.stack([cx, cy, cz], 1).to(dtype)
    pos = (2 * cell + 1 + rp) / R - 1
    c = torch.as_tensor(np.asarray(cam_center, dtype=np.float64)).to(dtype)
    return torch.linalg.norm(pos - c, dim=1)


def mse_loss(color, target):
    return ((color - target) ** 2).mean()


def sh_real_closed_form(v):
    """Independent degree-2 real SH table (e.g. Sloan, "Stupid SH tricks"),
    evaluated from the textbook normalisations.  Used to pin the constants
    and basis order of ``sh_irradiance``: the reference's basis is the real
    SH with the y axis as the pole, unsigned."""
    x, y, z = v[:, 0], v[:, 1], v[:, 2]
    k0 = 0.5 / math.sqrt(math.pi)
    k1 = math.sqrt(3.0 / (4.0 * math.pi))
    k2 = 0.5 * math.sqrt(15.0 / math.pi)
    k3 = 0.25 * math.sqrt(5.0 / math.pi)
    k4 = 0.25 * math.sqrt(15.0 / math.pi)
    return torch.stack([k0 * torch.ones_like(x), k1 * x, k1 * y, k1 * z, k2 * x * z,
                        k2 * x * y, k3 * (3 * y * y - (x * x + y * y + z * z)),
                        k2 * y * z, k4 * (z * z - x * x)], dim=1)
