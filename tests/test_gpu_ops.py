"""Per-op parity of the HIP op surface against the float64 oracle
(models/render_torch.py restated) on the small workload: 64^3 sphere grid,
100x100 image, 128 samples per ray, SH degree 1 in the 32-channel layout.

Tolerances (scale-aware, see helpers.scale_aware_close): forward values
1e-5, gradients 1e-4."""
import numpy as np
import pytest
import torch

from oracle import raygen
from oracle import render as orc
from qsdvr_amd import qsdvr, scenes
from qsdvr_amd import render as hr
from tests.helpers import assert_close

pytestmark = pytest.mark.gpu

FWD, BWD = 1e-5, 1e-4


@pytest.fixture(scope="module")
def wl():
    w = scenes.WORKLOADS["small"]
    cam = scenes.orbit_cameras(1, w.width, w.height)[0]
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    grid, data, gen = scenes.build_grids(w, seed=0)
    target = torch.rand((w.width * w.height, 3), generator=gen)
    dev = {k: torch.as_tensor(v).cuda() for k, v in smp.items() if isinstance(v, np.ndarray)}
    return {"w": w, "cam": cam, "smp": smp, "grid": grid, "data": data, "target": target,
            "dev": dev}


def _f64(t):
    return t.detach().double().cpu().requires_grad_(True)


def test_sample_sdf(wl):
    R = wl["w"].R
    d = wl["dev"]
    grid = wl["grid"].cuda()
    n = d["sdfPointList"].shape[0]
    out = torch.empty(n, device="cuda")
    qsdvr.SampleSDFForward(out, grid, d["sdfPointList"], d["sdfIndexList"], R)
    g64 = _f64(wl["grid"])
    ref = orc.sample_sdf(g64, torch.as_tensor(wl["smp"]["sdfPointList"]).double(),
                         torch.as_tensor(wl["smp"]["sdfIndexList"]), R)
    assert_close(out, ref, FWD, "SampleSDF forward")
    up = torch.randn(n, generator=torch.Generator().manual_seed(1))
    (ref * up.double()).sum().backward()
    gg = torch.zeros_like(grid)
    qsdvr.SampleSDFBackward(up.cuda(), grid, d["sdfPointList"], d["sdfIndexList"], gg, R)
    assert_close(gg.reshape(-1, 7), g64.grad.reshape(-1, 7), BWD, "SampleSDF backward")


def test_sample_normal(wl):
    d = wl["dev"]
    grid = wl["grid"].cuda()
    n = d["renderPointList"].shape[0]
    out = torch.empty((n, 3), device="cuda")
    qsdvr.SampleNormalForward(out, grid, d["renderPointList"], d["renderIndexList"])
    g64 = _f64(wl["grid"])
    ref = orc.sample_normal(g64, torch.as_tensor(wl["smp"]["renderPointList"]).double(),
                            torch.as_tensor(wl["smp"]["renderIndexList"]))
    assert_close(out, ref, FWD, "SampleNormal forward")
    up = torch.randn((n, 3), generator=torch.Generator().manual_seed(2))
    (ref * up.double()).sum().backward()
    gg = torch.zeros_like(grid)
    qsdvr.SampleNormalBackward(up.cuda(), grid, d["renderPointList"], d["renderIndexList"], gg)
    assert_close(gg.reshape(-1, 7), g64.grad.reshape(-1, 7), BWD, "SampleNormal backward")


def test_presorted_sample_normal_backward_matches_oracle(wl, monkeypatch):
    """SampleNormalBackward summed on GridInterpolationBackward's sort of the
    same samples (the reference's backward order) against the oracle's
    adjoint of SampleNormal (float64 autograd of render_torch.py:91-100)."""
    from qsdvr_amd import _lib
    d = wl["dev"]
    R = wl["w"].R
    grid, data = wl["grid"].cuda(), wl["data"].cuda()
    pts, idx = d["renderPointList"], d["renderIndexList"]
    n = pts.shape[0]
    g64 = _f64(wl["grid"])
    ref = orc.sample_normal(g64, torch.as_tensor(wl["smp"]["renderPointList"]).double(),
                            torch.as_tensor(wl["smp"]["renderIndexList"]))
    up = torch.randn((n, 3), generator=torch.Generator().manual_seed(12))
    (ref * up.double()).sum().backward()
    names = []
    real_call = _lib.call
    monkeypatch.setattr(_lib, "call", lambda name, *a: (names.append(name), real_call(name, *a)))
    qsdvr.GridInterpolationBackward(torch.zeros_like(data), pts, idx,
                                    torch.randn((n, data.shape[-1]), device="cuda"), R + 1)
    gg = torch.zeros_like(grid)
    qsdvr.SampleNormalBackward(up.cuda(), grid, pts, idx, gg)
    assert names[-1] == "qsdvr_sample_normal_backward_presorted"
    assert_close(gg.reshape(-1, 7), g64.grad.reshape(-1, 7), BWD, "presorted SampleNormal backward")


def test_grid_interpolation(wl):
    R = wl["w"].R
    d = wl["dev"]
    data = wl["data"].cuda()
    n = d["renderPointList"].shape[0]
    out = torch.empty((n, 32), device="cuda")
    qsdvr.GridInterpolationForward(data, d["renderPointList"], d["renderIndexList"], out, R + 1)
    d64 = _f64(wl["data"])
    ref = orc.grid_interpolation(d64, torch.as_tensor(wl["smp"]["renderPointList"]).double(),
                                 torch.as_tensor(wl["smp"]["renderIndexList"]), R)
    assert_close(out, ref, FWD, "GridInterpolation forward")
    up = torch.randn((n, 32), generator=torch.Generator().manual_seed(3))
    (ref * up.double()).sum().backward()
    gg = torch.zeros_like(data)
    qsdvr.GridInterpolationBackward(gg, d["renderPointList"], d["renderIndexList"], up.cuda(),
                                    R + 1)
    assert_close(gg, d64.grad, BWD, "GridInterpolation backward")


def test_shader(wl):
    g = torch.Generator().manual_seed(4)
    n = 50000
    nrm = torch.nn.functional.normalize(torch.randn((n, 3), generator=g), dim=1)
    view = torch.nn.functional.normalize(torch.randn((n, 3), generator=g), dim=1)
    data = torch.randn((n, 32), generator=g) * 0.3
    data[:, 27:] = torch.rand((n, 5), generator=g)
    out = torch.empty((n, 3), device="cuda")
    spec = torch.empty((n, 3), device="cuda")
    qsdvr.ShaderForward(out, nrm.cuda(), view.cuda(), data.cuda(), spec)
    n64, d64 = _f64(nrm), _f64(data)
    rgb, L = orc.shade(n64, view.double(), d64)
    assert_close(out, rgb, FWD, "Shader rgb")
    assert_close(spec, L, FWD, "Shader specular")
    up = torch.randn((n, 3), generator=g)
    (rgb * up.double()).sum().backward()
    gN = torch.empty((n, 3), device="cuda")
    gD = torch.empty((n, 32), device="cuda")
    qsdvr.ShaderBackward(up.cuda(), nrm.cuda(), view.cuda(), data.cuda(), spec, gN, gD)
    assert_close(gN, n64.grad, BWD, "Shader normal grad")
    assert_close(gD, d64.grad, BWD, "Shader data grad")


def _aggregate_inputs(wl):
    smp = wl["smp"]
    g = torch.Generator().manual_seed(5)
    rgb = torch.rand((smp["renderCount"], 3), generator=g)
    # a smooth sdf along each ray crossing zero, plus noise
    sdf = torch.randn(smp["sdfCount"], generator=g) * 0.05 + 0.3 - torch.rand(
        smp["sdfCount"], generator=g) * 0.6
    return rgb, sdf


@pytest.mark.parametrize("s", [5.0, 50.0])
def test_ray_aggregate(wl, s):
    smp = wl["smp"]
    rgb, sdf = _aggregate_inputs(wl)
    ray = wl["dev"]["rayList"]
    rays = ray.shape[0]
    out = torch.empty((rays, 3), device="cuda")
    alpha = torch.empty(rgb.shape[0], device="cuda")
    T = torch.empty(rgb.shape[0], device="cuda")
    qsdvr.RayAggregateForward(out, rgb.cuda(), sdf.cuda(), ray, alpha, T, s)
    r64, s64 = _f64(rgb), _f64(sdf)
    col, w = orc.composite(r64, s64, smp["rayList"], s, return_weights=True)
    assert_close(out, col, FWD, "RayAggregate colour")
    up = torch.randn((rays, 3), generator=torch.Generator().manual_seed(6))
    (col * up.double()).sum().backward()
    gR = torch.empty((rgb.shape[0], 3), device="cuda")
    gS = torch.empty(sdf.shape[0], device="cuda")
    qsdvr.RayAggregateBackward(up.cuda(), rgb.cuda(), sdf.cuda(), ray, alpha, T, gR, gS, s)
    assert_close(gR, r64.grad, BWD, "RayAggregate rgb grad")
    assert_close(gS, s64.grad, BWD, "RayAggregate sdf grad")


def test_layer_to_grid():
    R = 32
    x, y, z, off = orc.sphere_layers(R, 0.8, dtype=torch.float32)
    g = torch.Generator().manual_seed(7)
    x = x + 0.05 * torch.randn(R, generator=g)
    y = y + 0.05 * torch.randn(R, generator=g)
    z = z + 0.05 * torch.randn(R, generator=g)
    out = torch.empty((R, R, R, 7), device="cuda")
    qsdvr.LayerToGridForward(out, x, y, z, off)      # CPU layers, as the reference
    xs = [_f64(t) for t in (x, y, z, off)]
    ref = orc.layers_to_grid(*xs)
    assert_close(out.reshape(-1, 7), ref.reshape(-1, 7), FWD, "LayerToGrid forward")
    up = torch.randn((R, R, R, 7), generator=g)
    (ref * up.double()).sum().backward()
    gx, gy, gz, go = (torch.empty(R), torch.empty(R), torch.empty(R), torch.empty(4))
    qsdvr.LayerToGridBackward(up.cuda(), gx, gy, gz, go)
    for got, t, nm in ((gx, xs[0], "x"), (gy, xs[1], "y"), (gz, xs[2], "z"), (go, xs[3], "off")):
        assert_close(got, t.grad, BWD, f"LayerToGrid backward {nm}")
    # device layers give the same grid
    out2 = torch.empty_like(out)
    qsdvr.LayerToGridForward(out2, x.cuda(), y.cuda(), z.cuda(), off.cuda())
    assert torch.equal(out, out2)


def test_invalid_index_yields_nan_not_fault(wl):
    grid = wl["grid"].cuda()
    pts = torch.zeros((4, 3), device="cuda")
    idx = torch.tensor([0, -1, 64 ** 3, 5], dtype=torch.int32, device="cuda")
    out = torch.empty(4, device="cuda")
    qsdvr.SampleSDFForward(out, grid, pts, idx, 64)
    o = out.cpu()
    assert torch.isfinite(o[0]) and torch.isfinite(o[3])
    assert torch.isnan(o[1]) and torch.isnan(o[2])


def _three_scatters(smp_dev, grid, data, R, seed):
    """GridInterpolation, SampleSDF and SampleNormal backward of random
    upstream gradients on one sample list; returns the three grids."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    ns = smp_dev["sdfPointList"].shape[0]
    nr = smp_dev["renderPointList"].shape[0]
    up_s = torch.randn(ns, generator=g, device="cuda")
    up_n = torch.randn((nr, 3), generator=g, device="cuda")
    up_d = torch.randn((nr, data.shape[-1]), generator=g, device="cuda")
    gs = torch.zeros_like(grid)
    qsdvr.SampleSDFBackward(up_s, grid, smp_dev["sdfPointList"], smp_dev["sdfIndexList"], gs, R)
    gn = torch.zeros_like(grid)
    qsdvr.SampleNormalBackward(up_n, grid, smp_dev["renderPointList"],
                               smp_dev["renderIndexList"], gn)
    gd = torch.zeros_like(data)
    qsdvr.GridInterpolationBackward(gd, smp_dev["renderPointList"], smp_dev["renderIndexList"],
                                    up_d, R + 1)
    torch.cuda.synchronize()
    return gs, gn, gd


def test_sorted_scatters_match_atomics_on_a_full_pose(monkeypatch):
    """One 800x800 pose of the 128^3 torus (6.6e7 samples): the sorted-segment
    scatters equal the atomic kernels to 1e-4 and repeat bitwise."""
    from qsdvr_amd.camera import Camera
    w = scenes.WORKLOADS["plain"]
    cam = scenes.orbit_cameras(100, w.width, w.height)[17]
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    a = qsdvr.CameraArgs.from_camera(cam)
    cc = Camera(a, w.interval, w.R)
    cc.GenerateRayPoints(None)
    inp = cc.GetRayPoints()
    monkeypatch.setattr(qsdvr, "SCATTER", "sorted")
    s1 = _three_scatters(inp, grid, data, w.R, 5)
    s2 = _three_scatters(inp, grid, data, w.R, 5)
    monkeypatch.setattr(qsdvr, "SCATTER", "atomic")
    at = _three_scatters(inp, grid, data, w.R, 5)
    for name, x, y, z in zip(("SampleSDF", "SampleNormal", "GridInterpolation"), s1, s2, at):
        assert torch.equal(x, y), f"{name}: sorted scatter differs between two runs"
        rows = x if x.dim() == 2 else x.reshape(-1, 7)
        assert_close(rows, z.reshape(rows.shape), BWD, f"{name} sorted vs atomic")
        assert float(x.abs().sum()) > 0


@pytest.mark.parametrize("R", [9, 10])
def test_sorted_scatters_skip_invalid_indices_at_any_resolution(monkeypatch, R):
    """Random samples with out-of-range cell indices (skipped), at an odd
    resolution (the sorted GridInterpolation falls back to the atomic
    kernel) and an even one."""
    g = torch.Generator(device="cuda").manual_seed(8)
    n = 5000
    idx = torch.randint(-3, R ** 3 + 3, (n,), generator=g, device="cuda").to(torch.int32)
    pts = torch.rand((n, 3), generator=g, device="cuda") * 2 - 1
    grid = torch.randn((R, R, R, 7), generator=g, device="cuda")
    grid[..., 3] += 3.0  # keep the gradient norm away from zero
    data = torch.zeros(((R + 1) ** 3, 32), device="cuda")
    smp = {"sdfPointList": pts, "sdfIndexList": idx, "renderPointList": pts,
           "renderIndexList": idx}
    monkeypatch.setattr(qsdvr, "SCATTER", "sorted")
    s = _three_scatters(smp, grid, data, R, 9)
    monkeypatch.setattr(qsdvr, "SCATTER", "atomic")
    a = _three_scatters(smp, grid, data, R, 9)
    for x, y in zip(s, a):
        assert_close(x.reshape(-1, 7) if x.numel() % 7 == 0 and x.dim() == 4 else x,
                     y.reshape(-1, 7) if y.numel() % 7 == 0 and y.dim() == 4 else y, BWD,
                     "sorted vs atomic with invalid indices")


@pytest.mark.parametrize("C", [32, 53])
def test_sorted_interpolation_backward_matches_oracle(monkeypatch, C):
    """The sorted GridInterpolationBackward against the oracle's adjoint
    (float64 autograd of render_torch.py:128-143) on random samples that
    pile onto a few bricks, including the x = R rows that alias the next
    y-row's x = 0."""
    from oracle import render as orc
    R = 12
    g = torch.Generator().manual_seed(21)
    n = 20000
    cx = torch.randint(R - 5, R, (n,), generator=g)
    cy = torch.randint(0, R, (n,), generator=g)
    cz = torch.randint(3, 9, (n,), generator=g)
    idx = (cx + cy * R + cz * R * R).to(torch.int32)
    pts = torch.rand((n, 3), generator=g) * 2 - 1
    up = torch.randn((n, C), generator=g)
    d64 = torch.zeros(((R + 1) ** 3, C), dtype=torch.float64, requires_grad=True)
    out = orc.grid_interpolation(d64, pts.double(), idx, R)
    (out * up.double()).sum().backward()
    monkeypatch.setattr(qsdvr, "SCATTER", "sorted")
    gd = torch.zeros(((R + 1) ** 3, C), device="cuda")
    qsdvr.GridInterpolationBackward(gd, pts.cuda(), idx.cuda(), up.cuda(), R + 1)
    torch.cuda.synchronize()
    assert_close(gd, d64.grad, BWD, f"sorted GridInterpolationBackward (C={C})")
    if C == 32:
        # upstream rows and gradient grid 4 B off 16-B alignment: the per-channel
        # run kernel instead of the 16-B-quad one, same sums to tolerance
        up_m = torch.empty(n * C + 1, device="cuda")[1:].view(n, C)
        up_m.copy_(up.cuda())
        gd_m = torch.zeros(((R + 1) ** 3) * C + 1, device="cuda")[1:].view((R + 1) ** 3, C)
        assert up_m.data_ptr() % 16 != 0 and gd_m.data_ptr() % 16 != 0
        qsdvr.GridInterpolationBackward(gd_m, pts.cuda(), idx.cuda(), up_m, R + 1)
        torch.cuda.synchronize()
        assert_close(gd_m, d64.grad, BWD, "sorted GridInterpolationBackward, unaligned rows")


def test_normal_backward_reuses_the_interpolation_sort(monkeypatch):
    """On a full 800x800 pose: SampleNormalBackward right after
    GridInterpolationBackward on the same render indices sums its runs on
    that call's (parity class, cell) order instead of sorting again; it
    equals SampleNormalBackward's own sorted sums to 1e-4, repeats bitwise,
    and is not used when another call intervenes, the indices changed or the
    points are another tensor."""
    from qsdvr_amd import _lib
    from qsdvr_amd.camera import Camera
    w = scenes.WORKLOADS["plain"]
    cam = scenes.orbit_cameras(100, w.width, w.height)[23]
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    a = qsdvr.CameraArgs.from_camera(cam)
    cc = Camera(a, w.interval, w.R)
    cc.GenerateRayPoints(None)
    inp = cc.GetRayPoints()
    rp, ri = inp["renderPointList"], inp["renderIndexList"]
    g = torch.Generator(device="cuda").manual_seed(3)
    up_n = torch.randn((rp.shape[0], 3), generator=g, device="cuda")
    up_d = torch.randn((rp.shape[0], 32), generator=g, device="cuda")
    names = []
    real_call = _lib.call
    monkeypatch.setattr(_lib, "call", lambda name, *a: (names.append(name), real_call(name, *a)))
    monkeypatch.setattr(qsdvr, "SCATTER", "sorted")

    def normal():
        gn = torch.zeros_like(grid)
        qsdvr.SampleNormalBackward(up_n, grid, rp, ri, gn)
        return gn

    def interp():
        gd = torch.zeros_like(data)
        qsdvr.GridInterpolationBackward(gd, rp, ri, up_d, w.R + 1)

    own = normal()
    interp()
    reused = normal()
    interp()
    reused2 = normal()
    assert names.count("qsdvr_sample_normal_backward_presorted") == 2
    assert torch.equal(reused, reused2)
    assert_close(reused.reshape(-1, 7), own.reshape(-1, 7), BWD, "presorted vs own sort")
    names.clear()
    interp()
    qsdvr.SampleNormalForward(torch.empty_like(up_n), grid, rp, ri)   # an intervening call
    assert qsdvr._render_sort is None                                   # its workspace freed
    again = normal()
    assert "qsdvr_sample_normal_backward_presorted" not in names
    assert torch.equal(again, own)
    names.clear()
    interp()
    ri.add_(0)                                                        # in-place: new version
    assert torch.equal(normal(), own)
    assert "qsdvr_sample_normal_backward_presorted" not in names
    interp()
    gn = torch.zeros_like(grid)
    qsdvr.SampleNormalBackward(up_n, grid, rp.clone(), ri, gn)        # other point tensor
    assert torch.equal(gn, own)
    assert "qsdvr_sample_normal_backward_presorted" not in names


def test_sorted_interpolation_forward_equals_the_gather(monkeypatch):
    """GridInterpolationForward on its sort (one cell's run of samples at a
    time) gives the per-sample gather's rows bit for bit: on a full 800x800
    pose of the 128^3 torus, and on random samples with out-of-range indices
    (NaN rows).  An output 4 B off 16-B alignment takes the gather."""
    from qsdvr_amd import _lib
    from qsdvr_amd.camera import Camera
    from qsdvr_amd.qsdvr import _ptr, _stream
    names = []
    real_call = _lib.call
    monkeypatch.setattr(_lib, "call", lambda name, *a: (names.append(name), real_call(name, *a)))
    monkeypatch.setattr(qsdvr, "SCATTER", "sorted")
    w = scenes.WORKLOADS["plain"]
    cam = scenes.orbit_cameras(100, w.width, w.height)[31]
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cc = Camera(qsdvr.CameraArgs.from_camera(cam), w.interval, w.R)
    cc.GenerateRayPoints(None)
    inp = cc.GetRayPoints()
    rp, ri = inp["renderPointList"], inp["renderIndexList"]
    n = rp.shape[0]
    out = torch.empty((n, 32), device="cuda")
    qsdvr.GridInterpolationForward(data, rp, ri, out, w.R + 1)
    assert names[-1] == "qsdvr_grid_interpolation_forward_sorted"
    ref = torch.empty_like(out)
    real_call("qsdvr_grid_interpolation_forward", _ptr(data), _ptr(rp), _ptr(ri), _ptr(ref), n,
              w.R + 1, 32, _stream())
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    assert float(out.abs().sum()) > 0

    R = 10
    g = torch.Generator(device="cuda").manual_seed(12)
    m = 30011
    idx = torch.randint(-5, R ** 3 + 5, (m,), generator=g, device="cuda").to(torch.int32)
    pts = torch.rand((m, 3), generator=g, device="cuda") * 2 - 1
    dg = torch.randn(((R + 1) ** 3, 32), generator=g, device="cuda")
    o1 = torch.empty((m, 32), device="cuda")
    qsdvr.GridInterpolationForward(dg, pts, idx, o1, R + 1)
    assert names[-1] == "qsdvr_grid_interpolation_forward_sorted"
    o2 = torch.empty_like(o1)
    real_call("qsdvr_grid_interpolation_forward", _ptr(dg), _ptr(pts), _ptr(idx), _ptr(o2), m,
              R + 1, 32, _stream())
    torch.cuda.synchronize()
    bad = (idx < 0) | (idx >= R ** 3)
    assert int(bad.sum()) > 0
    assert bool(torch.isnan(o1[bad]).all()) and bool(torch.isfinite(o1[~bad]).all())
    assert torch.equal(o1[~bad], o2[~bad])
    o3 = torch.empty(m * 32 + 1, device="cuda")[1:].view(m, 32)
    qsdvr.GridInterpolationForward(dg, pts, idx, o3, R + 1)
    assert names[-1] == "qsdvr_grid_interpolation_forward"
    assert torch.equal(o3[~bad], o2[~bad])


def test_interpolation_backward_reuses_the_forward_sort(monkeypatch):
    """On a full 800x800 pose: GridInterpolationBackward on the point and index
    lists of the preceding sorted GridInterpolationForward sums on that
    forward's sort (no keys, no sort) and gives its own sort's sums bit for
    bit, even with other op calls in between; SampleNormalBackward then
    reuses the same sort.  An in-place change of the indices, other index
    tensors, or the point list being freed drops the kept sort."""
    import gc
    from qsdvr_amd import _lib
    from qsdvr_amd.camera import Camera
    names = []
    real_call = _lib.call
    monkeypatch.setattr(_lib, "call", lambda name, *a: (names.append(name), real_call(name, *a)))
    monkeypatch.setattr(qsdvr, "SCATTER", "sorted")
    w = scenes.WORKLOADS["plain"]
    cam = scenes.orbit_cameras(100, w.width, w.height)[41]
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cc = Camera(qsdvr.CameraArgs.from_camera(cam), w.interval, w.R)
    cc.GenerateRayPoints(None)
    inp = cc.GetRayPoints()
    rp, ri = inp["renderPointList"], inp["renderIndexList"]
    n = rp.shape[0]
    g = torch.Generator(device="cuda").manual_seed(4)
    up_d = torch.randn((n, 32), generator=g, device="cuda")
    up_n = torch.randn((n, 3), generator=g, device="cuda")

    def forward():
        qsdvr.GridInterpolationForward(data, rp, ri, torch.empty((n, 32), device="cuda"),
                                       w.R + 1)

    def backward(points=rp, index=ri):
        gd = torch.zeros_like(data)
        qsdvr.GridInterpolationBackward(gd, points, index, up_d, w.R + 1)
        return gd

    def normal():
        gn = torch.zeros_like(grid)
        qsdvr.SampleNormalBackward(up_n, grid, rp, ri, gn)
        return gn

    own = backward()
    own_n = normal()
    names.clear()
    forward()
    qsdvr.SampleNormalForward(torch.empty((n, 3), device="cuda"), grid, rp, ri)  # in between
    reused = backward()
    reused_n = normal()
    torch.cuda.synchronize()
    assert names.count("qsdvr_grid_interpolation_backward_presorted") == 1
    assert "qsdvr_grid_interpolation_backward_sorted" not in names
    assert names.count("qsdvr_sample_normal_backward_presorted") == 1
    assert qsdvr._forward_sort is None
    assert torch.equal(reused, own) and torch.equal(reused_n, own_n)
    names.clear()
    forward()
    ri.add_(0)                                    # in-place: new version
    assert torch.equal(backward(), own)
    assert "qsdvr_grid_interpolation_backward_presorted" not in names
    names.clear()
    forward()
    assert torch.equal(backward(index=ri.clone()), own)  # other index tensor
    assert "qsdvr_grid_interpolation_backward_presorted" not in names
    rp2, ri2 = rp.clone(), ri.clone()
    qsdvr.GridInterpolationForward(data, rp2, ri2, torch.empty((n, 32), device="cuda"), w.R + 1)
    assert qsdvr._forward_sort is not None
    del rp2
    gc.collect()
    qsdvr.SampleNormalForward(torch.empty((n, 3), device="cuda"), grid, rp, ri)
    assert qsdvr._forward_sort is None             # its point list freed: workspace dropped


def test_sorted_scatters_accept_an_unaligned_workspace():
    """The C ABI's sorted scatters take a caller workspace of the queried size
    at any address: its slices are placed at 256-B aligned addresses inside
    it, so a workspace 4 B off alignment gives the same bits as an aligned
    one (GridInterpolationBackward, the presorted SampleNormalBackward on its
    sort, SampleSDFBackward)."""
    import ctypes as C
    from qsdvr_amd import _lib
    from qsdvr_amd.qsdvr import _ptr, _stream
    R = 16
    g = torch.Generator(device="cuda").manual_seed(5)
    n = 50000
    idx = torch.randint(0, R ** 3, (n,), generator=g, device="cuda", dtype=torch.int32)
    pts = torch.rand((n, 3), generator=g, device="cuda")
    up = torch.randn((n, 32), generator=g, device="cuda")
    upn = torch.randn((n, 3), generator=g, device="cuda")
    ups = torch.randn((n,), generator=g, device="cuda")
    grid = torch.randn((R, R, R, 7), generator=g, device="cuda")
    nb = int(_lib.fn("qsdvr_sorted_scatter_workspace_size")(n, R ** 3))
    results = []
    for shift in (0, 4):
        raw = torch.empty(nb + 512, dtype=torch.uint8, device="cuda")
        ws = raw[(-raw.data_ptr()) % 256 + shift:][:nb]
        assert ws.data_ptr() % 256 == shift and ws.numel() == nb
        gd = torch.zeros(((R + 1) ** 3, 32), device="cuda")
        gn = torch.zeros_like(grid)
        gs = torch.zeros_like(grid)
        _lib.call("qsdvr_grid_interpolation_backward_sorted", _ptr(gd), _ptr(pts), _ptr(idx),
                  _ptr(up), n, R + 1, 32, _ptr(ws), nb, _stream())
        _lib.call("qsdvr_sample_normal_backward_presorted", _ptr(upn), _ptr(grid), _ptr(gn), n,
                  R, _ptr(ws), nb, _stream())
        _lib.call("qsdvr_sample_sdf_backward_sorted", _ptr(ups), _ptr(grid), _ptr(pts), _ptr(idx),
                  _ptr(gs), n, R ** 3, float(R), _ptr(ws), nb, _stream())
        torch.cuda.synchronize()
        results.append((gd, gn, gs))
    for a, b in zip(*results):
        assert torch.equal(a, b)
    assert results[0][0].abs().sum() > 0 and results[0][1].abs().sum() > 0
