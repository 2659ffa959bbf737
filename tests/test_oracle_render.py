"""Pins the render oracle (oracle/render.py, restating models/render_torch.py)
with checks that do not re-type its formulas: closed forms, special cases,
an independent SH table, literal-loop vs vectorised compositing, the NeuS
telescoping identity, the reference's own reverse recurrences for the
LayerToGrid adjoint, and float64 finite differences for every gradient."""
import math

import numpy as np
import pytest
import torch

from oracle import render as orc

D = torch.float64


def test_sphere_layers_closed_form():
    # init_as_sphere (render_torch.py:22-28): f = sum_axes (2i + 1 + p - R)^2 - (rR)^2
    R, r = 6, 0.7
    g = orc.layers_to_grid(*orc.sphere_layers(R, r))
    z, y, x = torch.meshgrid(torch.arange(R), torch.arange(R), torch.arange(R), indexing="ij")
    c = 1 - R
    for k, ax in ((0, x), (1, y), (2, z)):
        assert torch.all(g[..., k] == 1)
        assert torch.allclose(g[..., 3 + k], (2 * (2 * ax + c)).to(D))
    G = (2 * x + c) ** 2 + (2 * y + c) ** 2 + (2 * z + c) ** 2 - (r * R) ** 2
    assert torch.allclose(g[..., 6], G.to(D))


def _reference_layer_adjoint(grad):
    """The reverse recurrences of kernel/surface.cu:301-334 applied to the
    per-axis sums of the grid gradient (the reference's own algorithm)."""
    R = grad.shape[0]
    gx = [grad[..., 0].sum((0, 1)), grad[..., 3].sum((0, 1)), grad[..., 6].sum((0, 1))]
    gy = [grad[..., 1].sum((0, 2)), grad[..., 4].sum((0, 2)), grad[..., 6].sum((0, 2))]
    gz = [grad[..., 2].sum((1, 2)), grad[..., 5].sum((1, 2)), grad[..., 6].sum((1, 2))]
    out = []
    for a, d, c in (gx, gy, gz):
        a, d, c = a.clone(), d.clone(), c.clone()
        for i in range(R - 1, 0, -1):
            a[i] += c[i]
            a[i - 1] += 3 * c[i]
            d[i - 1] += 2 * c[i]
            c[i - 1] += c[i]
        for i in range(R - 1, 0, -1):
            a[i] += 2 * d[i]
            a[i - 1] += 2 * d[i]
            d[i - 1] += d[i]
        out.append((a, d[0]))
    off = torch.stack([out[0][1], out[1][1], out[2][1], grad[..., 6].sum()])
    return out[0][0], out[1][0], out[2][0], off


def test_layer_adjoint_matches_reference_recurrences_and_transpose():
    R = 7
    g = torch.Generator().manual_seed(0)
    ins = [torch.randn(R, generator=g, dtype=D).requires_grad_(True) for _ in range(3)]
    ins.append(torch.randn(4, generator=g, dtype=D).requires_grad_(True))
    out = orc.layers_to_grid(*ins)
    up = torch.randn(out.shape, generator=g, dtype=D)
    (out * up).sum().backward()
    ref = _reference_layer_adjoint(up)
    for a, b in zip([t.grad for t in ins], ref):
        assert torch.allclose(a, b, rtol=1e-12, atol=1e-9)
    # linear map: <J v, w> == <v, J^T w>
    v = [torch.randn(R, generator=g, dtype=D) for _ in range(3)] + [torch.randn(4, generator=g, dtype=D)]
    lhs = (orc.layers_to_grid(*v) * up).sum()
    rhs = sum((a * b).sum() for a, b in zip(v, ref))
    assert torch.allclose(lhs, rhs, rtol=1e-12)


def test_sdf_plane_and_sphere_normal():
    R = 8
    g = torch.Generator().manual_seed(1)
    n = 50
    grid = torch.zeros((R, R, R, 7), dtype=D)
    grid[..., 3:6] = torch.tensor([0.3, -0.4, 1.2], dtype=D)
    grid[..., 6] = 0.25
    pts = torch.rand((n, 3), generator=g, dtype=D) * 2 - 1
    idx = torch.randint(0, R ** 3, (n,), generator=g)
    sdf = orc.sample_sdf(grid, pts, idx, R)
    exp = (pts @ torch.tensor([0.3, -0.4, 1.2], dtype=D) + 0.25) / math.sqrt(0.09 + 0.16 + 1.44) / R
    assert torch.allclose(sdf, exp, rtol=1e-13)
    nrm = orc.sample_normal(grid, pts, idx)
    assert torch.allclose(nrm, torch.tensor([0.3, -0.4, 1.2], dtype=D).expand(n, 3) / math.sqrt(1.69))
    # sphere quadric: the normal is the radial direction
    sph = orc.layers_to_grid(*orc.sphere_layers(R, 0.5))
    cell = torch.randint(0, R, (n, 3), generator=g)
    idx = cell[:, 0] + cell[:, 1] * R + cell[:, 2] * R * R
    u = 2 * cell + 1 + pts - R                     # position relative to the centre
    nrm = orc.sample_normal(sph, pts, idx)
    assert torch.allclose(nrm, u / u.norm(dim=1, keepdim=True), atol=1e-12)
    # and f/|grad f| = (rho^2 - r^2) / (2 rho) in cell units
    rho = u.norm(dim=1)
    exp = (rho ** 2 - (0.5 * R) ** 2) / (2 * rho) / R
    assert torch.allclose(orc.sample_sdf(sph, pts, idx, R), exp, atol=1e-12)


def test_interpolation_weights_and_transpose():
    R, C = 5, 4
    g = torch.Generator().manual_seed(2)
    n = 40
    pts = torch.rand((n, 3), generator=g, dtype=D) * 2 - 1
    idx = torch.randint(0, R ** 3, (n,), generator=g)
    const = torch.full(((R + 1) ** 3, C), 0.7, dtype=D)
    assert torch.allclose(orc.grid_interpolation(const, pts, idx, R), torch.full((n, C), 0.7, dtype=D))
    # indicator rows: output = the weight of that row; rows i + dx + dy R + dz R^2,
    # weights prod(dx ? p : 1 - p) with the raw local point
    for dx, dy, dz in ((0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)):
        data = torch.zeros(((R + 1) ** 3, 1), dtype=D)
        i0 = 2 + 1 * R + 2 * R * R
        data[i0 + dx + dy * R + dz * R * R] = 1
        q = torch.zeros(1, dtype=torch.long) + i0
        p = pts[:1]
        wexp = ((p[0, 0] if dx else 1 - p[0, 0]) * (p[0, 1] if dy else 1 - p[0, 1])
                * (p[0, 2] if dz else 1 - p[0, 2]))
        assert torch.allclose(orc.grid_interpolation(data, p, q, R)[0, 0], wexp)
    data = torch.randn(((R + 1) ** 3, C), generator=g, dtype=D).requires_grad_(True)
    out = orc.grid_interpolation(data, pts, idx, R)
    up = torch.randn(out.shape, generator=g, dtype=D)
    (out * up).sum().backward()
    v = torch.randn(data.shape, generator=g, dtype=D)
    assert torch.allclose((orc.grid_interpolation(v, pts, idx, R) * up).sum(), (v * data.grad).sum())


def test_sh_matches_independent_table():
    g = torch.Generator().manual_seed(3)
    v = torch.nn.functional.normalize(torch.randn((30, 3), generator=g, dtype=D), dim=1)
    table = orc.sh_real_closed_form(v)
    K = torch.as_tensor(orc.SH_FACTOR_F32, dtype=D)
    for k in range(9):
        coef = torch.zeros((30, 9, 3), dtype=D)
        coef[:, k, :] = 1.0
        got = orc.sh_irradiance(v, coef, factors=K.numpy())[:, 0]
        assert torch.allclose(got, table[:, k], rtol=1e-7, atol=1e-8), k  # float32 constants
    # the float32 constants are the float32 roundings of the closed forms
    exact = [0.5 / math.sqrt(math.pi), math.sqrt(3 / (4 * math.pi)), math.sqrt(15 / (4 * math.pi)),
             math.sqrt(5 / (16 * math.pi)), math.sqrt(15 / (16 * math.pi))]
    assert [float(np.float32(x)) for x in exact] == [float(orc.SH_FACTOR_F32[i]) for i in (0, 1, 4, 6, 8)]


def _sphere_quadrature(n=12, m=24):
    """Gauss-Legendre in y (the pole) times uniform azimuth: exact for the
    products of two degree-3 polynomials on the unit sphere."""
    ys, wy = np.polynomial.legendre.leggauss(n)
    ph = np.arange(m) * 2 * np.pi / m
    yy, pp = np.meshgrid(ys, ph, indexing="ij")
    st = np.sqrt(1 - yy ** 2)
    v = np.stack([(st * np.cos(pp)).ravel(), yy.ravel(), (st * np.sin(pp)).ravel()], 1)
    w = np.outer(wy, np.full(m, 2 * np.pi / m)).ravel()
    return torch.as_tensor(v, dtype=D), torch.as_tensor(w, dtype=D)


def test_sh_degree3_orthonormal_harmonic_and_textbook():
    """Degree 3 (no reference counterpart) is pinned three ways: the 16
    functions with the float32 constants are orthonormal on the sphere; every
    degree-3 polynomial is harmonic; and each equals the textbook z-pole
    real SH of the same m (Sloan, "Stupid SH tricks") under the axis
    substitution that maps the textbook's degree-2 table onto
    render_torch's: P(x, y, z) = Q(z, x, y)."""
    v, w = _sphere_quadrature()
    K = torch.as_tensor(orc.SH_FACTOR16_F32, dtype=D)
    Y = orc.sh_basis(v, 16) * K
    G = Y.t() @ (Y * w[:, None])
    assert torch.allclose(G, torch.eye(16, dtype=D), atol=1e-6)
    exact = [math.sqrt(35 / (32 * math.pi)), math.sqrt(105 / (4 * math.pi)),
             math.sqrt(21 / (32 * math.pi)), math.sqrt(7 / (16 * math.pi)),
             math.sqrt(21 / (32 * math.pi)), math.sqrt(105 / (16 * math.pi)),
             math.sqrt(35 / (32 * math.pi))]
    assert [float(np.float32(x)) for x in exact] == [float(x) for x in orc.SH3_FACTOR_F32]
    Ye = orc.sh_basis(v, 16)[:, 9:] * torch.as_tensor(exact, dtype=D)
    Ge = Ye.t() @ (Ye * w[:, None])
    assert torch.allclose(Ge, torch.eye(7, dtype=D), atol=1e-12)
    # harmonic: the Laplacian of each homogeneous cubic vanishes
    g = torch.Generator().manual_seed(12)
    p = torch.randn((7, 3), generator=g, dtype=D).requires_grad_(True)
    for k in range(9, 16):
        f = lambda q: orc.sh_basis(q.reshape(1, 3), 16)[0, k]
        for i in range(7):
            H = torch.autograd.functional.hessian(f, p[i].detach())
            assert abs(float(torch.trace(H))) < 1e-12, k
    # textbook z-pole degree-3 table, m = -3..3, evaluated at (z, x, y)
    x, y, z = v[:, 2], v[:, 0], v[:, 1]
    Q = torch.stack([y * (3 * x * x - y * y), x * y * z, y * (4 * z * z - x * x - y * y),
                     z * (2 * z * z - 3 * x * x - 3 * y * y), x * (4 * z * z - x * x - y * y),
                     z * (x * x - y * y), x * (x * x - 3 * y * y)], 1)
    assert torch.allclose(orc.sh_basis(v, 16)[:, 9:], Q, atol=1e-13)
    # the same substitution maps the textbook degree <= 2 table onto render_torch's
    Q2 = torch.stack([torch.ones_like(x), y, z, x, x * y, y * z, 2 * z * z - x * x - y * y,
                      x * z, x * x - y * y], 1)
    assert torch.allclose(orc.sh_basis(v, 9), Q2, atol=1e-13)


def test_shade_degree3_extends_degree2():
    """The 53-channel shading with all degree-3 coefficients zero equals the
    reference's 32-channel shading of the same other channels."""
    g = torch.Generator().manual_seed(13)
    n = 25
    nrm = torch.nn.functional.normalize(torch.randn((n, 3), generator=g, dtype=D), dim=1)
    view = torch.nn.functional.normalize(torch.randn((n, 3), generator=g, dtype=D), dim=1)
    d32 = torch.randn((n, 32), generator=g, dtype=D)
    d53 = torch.zeros((n, 53), dtype=D)
    d53[:, :27] = d32[:, :27]
    d53[:, 48:53] = d32[:, 27:32]
    rgb32, L32 = orc.shade(nrm, view, d32)
    rgb53, L53 = orc.shade(nrm, view, d53)
    assert torch.allclose(rgb32, rgb53, atol=1e-14) and torch.allclose(L32, L53, atol=1e-14)
    # and a single degree-3 coefficient adds exactly its basis term
    d53[:, 3 * 12 + 1] = 1.0  # coefficient 12, green
    r = view - 2 * (view * nrm).sum(1, keepdim=True) * nrm
    _, L = orc.shade(nrm, view, d53)
    y = r[:, 1]
    term = float(orc.SH3_FACTOR_F32[3]) * y * (2 * y * y - 3 * r[:, 0] ** 2 - 3 * r[:, 2] ** 2)
    assert torch.allclose(L[:, 1] - L32[:, 1], term, atol=1e-12)


def test_shading_special_cases():
    g = torch.Generator().manual_seed(4)
    n = 20
    nrm = torch.nn.functional.normalize(torch.randn((n, 3), generator=g, dtype=D), dim=1)
    view = torch.nn.functional.normalize(torch.randn((n, 3), generator=g, dtype=D), dim=1)
    data = torch.randn((n, 32), generator=g, dtype=D)
    data[:, 30] = 0.0
    rgb, _ = orc.shade(nrm, view, data)
    assert torch.allclose(rgb, data[:, 27:30] * data[:, 31:32])
    data = torch.zeros((n, 32), dtype=D)
    data[:, 0:3] = torch.randn((n, 3), generator=g, dtype=D)
    data[:, 30] = 1.0
    data[:, 31] = 2.0
    rgb, L = orc.shade(nrm, view, data)
    assert torch.allclose(rgb, 2.0 * float(orc.SH_FACTOR_F32[0]) * data[:, 0:3])
    # reflection: r = v - 2 (v.n) n keeps |r| = 1 and is mirror-symmetric
    data = torch.zeros((n, 32), dtype=D)
    data[:, 3 * 1 + 0] = 1.0  # coefficient 1 (x), red
    data[:, 30] = 1.0
    data[:, 31] = 1.0
    rgb, _ = orc.shade(nrm, view, data)
    r = view - 2 * (view * nrm).sum(1, keepdim=True) * nrm
    assert torch.allclose(rgb[:, 0], float(orc.SH_FACTOR_F32[1]) * r[:, 0])


def _rays(counts):
    counts = np.asarray(counts, np.int64)
    ren = np.maximum(counts - 1, 0)
    ray = np.stack([counts, np.cumsum(counts) - counts, ren, np.cumsum(ren) - ren], 1)
    return ray.astype(np.int32)


def test_composite_literal_vs_vectorised_and_invariants():
    g = torch.Generator().manual_seed(5)
    counts = [0, 2, 5, 9, 0, 3, 12]
    ray = _rays(counts)
    S, Rn = sum(counts), sum(max(c - 1, 0) for c in counts)
    sdf = torch.randn(S, generator=g, dtype=D) * 0.3
    rgb = torch.rand((Rn, 3), generator=g, dtype=D)
    for s in (1.0, 5.0, 40.0):
        a = orc.composite(rgb, sdf, ray, s)
        b = orc.composite_literal(rgb, sdf, ray, s)
        assert torch.allclose(a, b, rtol=1e-13, atol=1e-15)
        _, w = orc.composite(rgb, sdf, ray, s, return_weights=True)
        assert torch.all(w >= 0) and torch.all(w.sum(1) <= 1 + 1e-12)
    assert torch.all(orc.composite(rgb, sdf, ray, 5.0)[[0, 4]] == 0)


def test_composite_neus_telescoping():
    # monotonically decreasing sdf: no clamp, T_k = Phi(sdf_k) / Phi(sdf_0),
    # opacity = 1 - Phi(sdf_n) / Phi(sdf_0), Phi(x) = sigmoid(s x)
    s = 7.0
    sdf = torch.linspace(0.8, -0.6, 15, dtype=D)
    ray = _rays([15])
    rgb = torch.ones((14, 3), dtype=D)
    col, w = orc.composite(rgb, sdf, ray, s, return_weights=True)
    phi = torch.sigmoid(s * sdf)
    assert torch.allclose(col[0, 0], 1 - phi[-1] / phi[0], rtol=1e-13)
    T = 1 - torch.cat([torch.zeros(1, dtype=D), torch.cumsum(w[0], 0)[:-1]])  # T_k = 1 - sum_{i<k} w_i
    assert torch.allclose(phi[:-1] / phi[0], T, rtol=1e-12)
    # single step closed form
    sdf2 = torch.tensor([0.1, -0.2], dtype=D)
    c2 = orc.composite(torch.tensor([[0.2, 0.5, 0.9]], dtype=D), sdf2, _rays([2]), s)
    e0, e1 = math.exp(-s * 0.1), math.exp(s * 0.2)
    a = (e1 - e0) / (1 + e1)
    assert torch.allclose(c2[0], a * torch.tensor([0.2, 0.5, 0.9], dtype=D))


def _axis_camera(cx_world, W=5, near=0.4, far=4.0):
    """A camera at (cx_world, 0, 0) looking along +x; the centre pixel of an
    odd-sized image maps exactly onto the optical axis (u + cx = 0)."""
    from qsdvr_amd import scenes
    cam = scenes.look_at_camera(np.array([cx_world, 0.0, 0.0]), W, W, near=near, far=far)
    rot = cam.rotation.reshape(3, 3)
    assert np.array_equal(rot[:, 2], np.array([1, 0, 0], np.float32))
    return cam


def _plane_grid(R, x0):
    """Quadric grid whose SDF is the plane x = x0 (positive before it along
    +x): A = B = C = 0, D = -1, G = R (x0 - cell centre x), so
    f / |grad f| / R = x0 - x exactly."""
    g = torch.zeros((R, R, R, 7), dtype=D)
    xc = (2 * torch.arange(R, dtype=D) + 1) / R - 1
    g[..., 3] = -1.0
    g[..., 6] = (R * (x0 - xc)).view(1, 1, R)
    return g


@pytest.mark.parametrize("s", [7.0, 300.0])
def test_depth_closed_form_planar_sdf(s):
    """Depth D = sum_i T_i alpha_i t_i with t_i the distance from the camera
    centre to render sample i, on an axis-aligned ray through a plane.  The
    SDF falls monotonically along the ray, so alpha never clamps and the
    weights telescope: T_k = (1 + e_0) / (1 + e_k), w_k = (1 + e_0)
    (sig_k - sig_{k+1}) with sig_k = 1 / (1 + e_k).  Positions, SDF values
    and distances are rebuilt here in float64 from the camera geometry, not
    from the oracle's lists.  At large s the depth is the camera-to-plane
    distance to within one step.  Measuring t from the in-cube ray origin
    instead of the camera centre moves D by about 1.6 and fails both."""
    from oracle import raygen
    R, N, x0, c0 = 16, 64, 0.13, -2.6
    cam = _axis_camera(c0)
    interval = np.float32(2 * math.sqrt(3) / N)
    pix = np.array([12])                                    # centre pixel of 5x5
    smp = raygen.ray_samples(cam.as_dict(), interval, R, pixels=pix)
    grid = _plane_grid(R, x0)
    data = torch.zeros(((R + 1) ** 3, 32), dtype=D)
    data[:, 27:30] = 0.5                                    # diffuse
    data[:, 31] = 1.0                                       # ao, metallic 0
    centre = cam.translation.astype(np.float64)
    col, dep = orc.render(grid, data, smp, R, s, with_depth=True, cam_center=centre)
    # float64 geometry: the ray is x(t) = c0 + t along +x from the camera centre
    dt = float(interval)
    cx = float(np.float32(c0))
    xn = cx + float(np.float32(cam.near))                   # near point, depth = distance
    j = np.arange(0, 400)
    xj = xn + j * dt
    inside = np.abs(xj) < 0.9999
    j0, cnt = int(np.argmax(inside)), int(inside.sum())
    assert int(smp["rayList"][0, 0]) == cnt
    ori = xj[j0]
    xs = ori + np.arange(cnt) * dt                          # sdf samples
    tr = (ori - cx) + (np.arange(cnt - 1) + 0.5) * dt       # render-sample distances
    e = np.exp(-s * (x0 - xs))
    sig = 1.0 / (1.0 + e)
    w = (1.0 + e[0]) * (sig[:-1] - sig[1:])
    assert np.all(w >= 0)
    d_closed = float((w * tr).sum())
    assert abs(float(dep[0]) - d_closed) <= 1e-5 * d_closed
    assert abs(float(col[0, 0]) - 0.5 * w.sum()) <= 1e-6
    if s >= 100:
        assert abs(d_closed - (x0 - cx)) <= dt
        assert abs(float(dep[0]) - (x0 - cx)) <= dt
    # the mistaken distance (from the in-cube origin) is far off
    assert abs(float((w * (tr - (ori - cx))).sum()) - float(dep[0])) > 1.0


def test_clamp_gradient_passes_at_zero():
    # equal consecutive sdf -> pre-clamp alpha exactly 0; torch.clamp passes
    # the gradient there (the reference CUDA kernel zeroes it)
    sdf = torch.tensor([0.2, 0.2, -0.1], dtype=D, requires_grad=True)
    rgb = torch.tensor([[1.0, 1.0, 1.0], [0.5, 0.5, 0.5]], dtype=D)
    col = orc.composite(rgb, sdf, _rays([3]), 5.0)
    col.sum().backward()
    assert sdf.grad[1].abs() > 0 and sdf.grad[0].abs() > 0


@pytest.mark.parametrize("fn", ["sdf", "normal", "interp", "shade", "shade53", "composite"])
def test_finite_differences(fn):
    g = torch.Generator().manual_seed(6)
    R = 3
    if fn in ("sdf", "normal"):
        grid = (torch.randn((R, R, R, 7), generator=g, dtype=D) + torch.tensor(
            [0, 0, 0, 2, 2, 2, 0], dtype=D)).requires_grad_(True)
        pts = torch.rand((6, 3), generator=g, dtype=D) * 2 - 1
        idx = torch.randint(0, R ** 3, (6,), generator=g)
        f = (lambda q: orc.sample_sdf(q, pts, idx, R)) if fn == "sdf" else (
            lambda q: orc.sample_normal(q, pts, idx))
        assert torch.autograd.gradcheck(f, (grid,))
    elif fn == "interp":
        data = torch.randn(((R + 1) ** 3, 4), generator=g, dtype=D).requires_grad_(True)
        pts = torch.rand((6, 3), generator=g, dtype=D) * 2 - 1
        idx = torch.randint(0, R ** 3, (6,), generator=g)
        assert torch.autograd.gradcheck(lambda d: orc.grid_interpolation(d, pts, idx, R), (data,))
    elif fn in ("shade", "shade53"):
        nrm = torch.nn.functional.normalize(torch.randn((5, 3), generator=g, dtype=D), dim=1)
        view = torch.nn.functional.normalize(torch.randn((5, 3), generator=g, dtype=D), dim=1)
        data = torch.randn((5, 32 if fn == "shade" else 53), generator=g, dtype=D)
        assert torch.autograd.gradcheck(lambda n, d: orc.shade(n, view, d)[0],
                                        (nrm.requires_grad_(True), data.requires_grad_(True)))
    else:
        ray = _rays([4, 0, 6])
        sdf = (torch.randn(10, generator=g, dtype=D) * 0.3).requires_grad_(True)
        rgb = torch.rand((8, 3), generator=g, dtype=D).requires_grad_(True)
        assert torch.autograd.gradcheck(lambda r, s: orc.composite(r, s, ray, 5.0), (rgb, sdf))


def test_compacted_grids_give_the_full_grid_result():
    """The oracle on the rows a few rays touch (compacted grids, rows looked
    up through ``compact_index``) equals the oracle on the full grids,
    bit for bit, and so do the gradients at those rows; the pixel-subset ray
    generator gives the full frame's rows for those pixels."""
    from oracle import raygen
    from qsdvr_amd import scenes
    w = scenes.Workload("t", 12, 2, 16, 12, 40, 1, "torus", 0.005)
    cam = scenes.orbit_cameras(3, w.width, w.height)[1]
    grid, data, gen = scenes.build_grids(w, seed=3)
    pix = np.arange(5, w.width * w.height, 7)
    full = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    sub = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R, pixels=pix)
    assert np.array_equal(sub["rayList"][:, [0, 2]], full["rayList"][pix][:, [0, 2]])
    centre = cam.translation.astype(np.float64)
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    c_full, d_full = orc.render(g64, d64, full, w.R, w.s, with_depth=True, cam_center=centre)
    (c_full[pix].sum() + d_full[pix].sum()).backward()
    cells = np.unique(np.concatenate([sub["sdfIndexList"], sub["renderIndexList"]]))
    verts = np.unique(orc.vertex_rows(sub["renderIndexList"], w.R).numpy())
    gc = grid.double().reshape(-1, 7)[cells].clone().requires_grad_(True)
    dc = data.double()[verts].clone().requires_grad_(True)
    c_sub, d_sub = orc.render(gc, dc, sub, w.R, w.s, with_depth=True, cam_center=centre,
                              cell_rows=cells, vertex_rows=verts)
    assert torch.equal(c_sub, c_full[pix].detach()) and torch.equal(d_sub, d_full[pix].detach())
    (c_sub.sum() + d_sub.sum()).backward()
    assert torch.allclose(gc.grad, g64.grad.reshape(-1, 7)[cells], rtol=1e-12, atol=1e-15)
    assert torch.allclose(dc.grad, d64.grad[verts], rtol=1e-12, atol=1e-15)
    with pytest.raises(KeyError):
        orc.compact_index(verts[1:], torch.as_tensor(verts[:1]))
