"""End-to-end parity: the op-surface composition (models/render.py call
sequence on the HIP ops) and the fused renderer against the float64 oracle
(render_torch restated) on the small workload, and the fused renderer
against the op surface on one full 800x800 pose of the 128^3 torus."""
import numpy as np
import pytest
import torch

from oracle import raygen
from oracle import render as orc
from qsdvr_amd import fused, qsdvr, scenes
from qsdvr_amd import render as hr
from qsdvr_amd.camera import Camera
from tests.helpers import assert_close

pytestmark = pytest.mark.gpu

FWD, BWD = 1e-5, 1e-4


def _setup(w, pose=0, n_poses=1, seed=0):
    cam = scenes.orbit_cameras(max(n_poses, 1), w.width, w.height)[pose]
    grid, data, gen = scenes.build_grids(w, seed=seed)
    target = torch.rand((w.width * w.height, 3), generator=gen)
    return cam, grid, data, target


def _oracle(w, cam, grid, data, target, s):
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    col, dep = orc.render(g64, d64, smp, w.R, s, with_depth=True,
                          cam_center=cam.translation.astype(np.float64))
    loss = orc.mse_loss(col, target.double()) + 0.01 * dep.mean()
    loss.backward()
    return col.detach(), dep.detach(), g64.grad, d64.grad


@pytest.fixture(scope="module")
def small():
    w = scenes.WORKLOADS["small"]
    cam, grid, data, target = _setup(w)
    ref = _oracle(w, cam, grid, data, target, w.s)
    return w, cam, grid, data, target, ref


def test_op_surface_matches_oracle(small):
    w, cam, grid, data, target, (col, dep, gq, gd) = small
    a = qsdvr.CameraArgs.from_camera(cam)
    c = Camera(a, w.interval, w.R)
    c.GenerateRayPoints(None)
    inp = c.GetRayPoints()
    G = grid.cuda().requires_grad_(True)
    D = data.cuda().requires_grad_(True)
    out = hr.render_lists(G, D, inp, w.s)
    assert_close(out, col, FWD, "op-surface colour")
    loss = ((out - target.cuda()) ** 2).mean()
    loss.backward()
    # the oracle loss also has a depth term; rebuild the colour-only gradient
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    c64 = orc.render(g64, d64, smp, w.R, w.s)
    orc.mse_loss(c64, target.double()).backward()
    assert_close(G.grad.reshape(-1, 7), g64.grad.reshape(-1, 7), BWD, "op-surface sdf-grid grad")
    assert_close(D.grad, d64.grad, BWD, "op-surface appearance grad")


@pytest.mark.parametrize("mode", [fused.SCATTER_TILE, fused.SCATTER_ATOMIC])
def test_fused_matches_oracle(small, mode):
    w, cam, grid, data, target, (col, dep, gq, gd) = small
    G = grid.cuda().requires_grad_(True)
    D = data.cuda().requires_grad_(True)
    c, d = fused.render(G, D, [cam], w.interval, w.s, scatter_mode=mode)
    assert_close(c, col, FWD, "fused colour")
    assert_close(d, dep, FWD, "fused depth")
    loss = ((c - target.cuda()) ** 2).mean() + 0.01 * d.mean()
    loss.backward()
    assert_close(G.grad.reshape(-1, 7), gq.reshape(-1, 7), BWD, "fused sdf-grid grad")
    assert_close(D.grad, gd, BWD, "fused appearance grad")


def test_fused_matches_op_surface_full_pose():
    w = scenes.WORKLOADS["plain"]
    cam, grid, data, target = _setup(w, pose=17, n_poses=100)
    G = grid.cuda().requires_grad_(True)
    D = data.cuda().requires_grad_(True)
    c, _ = fused.render(G, D, [cam], w.interval, w.s)
    ((c - target.cuda()) ** 2).mean().backward()
    a = qsdvr.CameraArgs.from_camera(cam)
    cc = Camera(a, w.interval, w.R)
    cc.GenerateRayPoints(None)
    inp = cc.GetRayPoints()
    G2 = grid.cuda().requires_grad_(True)
    D2 = data.cuda().requires_grad_(True)
    out = hr.render_lists(G2, D2, inp, w.s)
    ((out - target.cuda()) ** 2).mean().backward()
    assert_close(c, out, FWD, "fused vs op colour")
    assert_close(G.grad.reshape(-1, 7), G2.grad.reshape(-1, 7), BWD, "fused vs op sdf grad")
    assert_close(D.grad, D2.grad, BWD, "fused vs op appearance grad")


def test_fused_ray_list_matches_full_frame(small):
    w, cam, grid, data, target, _ = small
    G = grid.cuda()
    D = data.cuda()
    c_full, d_full = fused.render(G, D, [cam], w.interval, w.s)
    g = torch.Generator().manual_seed(11)
    pick = torch.randperm(w.width * w.height, generator=g)[:3000].to(torch.int32)
    c_sub, d_sub = fused.render(G, D, [cam], w.interval, w.s, ray_pixel=pick.cuda())
    assert torch.equal(c_sub.cpu(), c_full.cpu()[pick.long()])
    assert torch.equal(d_sub.cpu(), d_full.cpu()[pick.long()])


def test_degenerate_cells_keep_gradients_finite(small):
    """A cell whose quadric has no gradient (A..F = 0) has an infinite SDF
    (f / |grad f| with |grad f| = 0).  The renderers give its sdf samples
    alpha 0, so their upstream gradient is 0, and a zero upstream must add
    nothing rather than 0 * inf = NaN: every gradient stays finite, in every
    scatter mode and on the op surface."""
    w, cam, grid, data, target, _ = small
    g = grid.clone()
    g[w.R // 2 - 1:w.R // 2 + 1, :, :, :6] = 0.0  # a slab of degenerate cells across the object
    g[..., 6] = torch.where(g[..., 6] == 0, torch.full_like(g[..., 6], 0.5), g[..., 6])
    tgt = target.cuda()
    for mode in (fused.SCATTER_TILE, fused.SCATTER_ATOMIC, fused.SCATTER_DET):
        G = g.cuda().requires_grad_(True)
        D = data.cuda().requires_grad_(True)
        c, d = fused.render(G, D, [cam], w.interval, w.s, scatter_mode=mode)
        ((c - tgt) ** 2).mean().backward()
        assert torch.isfinite(c).all() and torch.isfinite(d).all(), mode
        assert torch.isfinite(G.grad).all() and torch.isfinite(D.grad).all(), mode
    a = qsdvr.CameraArgs.from_camera(cam)
    cc = Camera(a, w.interval, w.R)
    cc.GenerateRayPoints(None)
    G = g.cuda().requires_grad_(True)
    D = data.cuda().requires_grad_(True)
    out = hr.render_lists(G, D, cc.GetRayPoints(), w.s)
    ((out - tgt) ** 2).mean().backward()
    assert torch.isfinite(out).all()
    assert torch.isfinite(G.grad).all() and torch.isfinite(D.grad).all()


def test_cost_ordered_backward_matches_launch_order(monkeypatch):
    """The tile adjoint run heaviest-first by the forward's per-tile cost
    (qsdvr_fused_backward_ordered, the default for whole frames) gives the
    launch-order adjoint's gradients (float atomics in another order: to
    1e-4) and the same colours bit for bit, on a full 800x800 pose and on
    two poses in one call; every tile gets a cost."""
    from qsdvr_amd import _lib
    from qsdvr_amd.qsdvr import _ptr, _stream
    w = scenes.WORKLOADS["plain"]
    cams = scenes.orbit_cameras(100, w.width, w.height)
    grid, data, _ = scenes.build_grids(w, seed=0)
    tg = torch.rand((2 * w.width * w.height, 3), generator=torch.Generator().manual_seed(4)).cuda()
    for pick in ([cams[11]], [cams[7], cams[30]]):
        out = {}
        for order in (False, True):
            monkeypatch.setattr(fused, "TILE_ORDER", order)
            G = grid.cuda().requires_grad_(True)
            D = data.cuda().requires_grad_(True)
            c, _ = fused.render(G, D, pick, w.interval, w.s)
            ((c - tg[:c.shape[0]]) ** 2).mean().backward()
            out[order] = (c.detach(), G.grad, D.grad)
        assert torch.equal(out[False][0], out[True][0])
        assert_close(out[True][1].reshape(-1, 7), out[False][1].reshape(-1, 7), BWD,
                     "ordered vs launch-order sdf grad")
        assert_close(out[True][2], out[False][2], BWD, "ordered vs launch-order appearance grad")
    # the costs themselves: every tile written, the object's tiles costlier than empty ones
    cb = fused.camera_buffer(cams[11:12])
    n_tiles = ((w.width + 7) // 8) * ((w.height + 7) // 8)
    cost = torch.full((n_tiles,), -1, dtype=torch.int32, device="cuda")
    col = torch.empty((w.width * w.height, 3), device="cuda")
    dep = torch.empty((w.width * w.height,), device="cuda")
    G, D = grid.cuda(), data.cuda()
    _lib.call("qsdvr_fused_forward_costed", _ptr(cb), 1, w.width, w.height, None, None, 0, _ptr(G),
              _ptr(D), w.R, 32, float(w.interval), float(w.s), _ptr(col), _ptr(dep), _ptr(cost),
              _stream())
    torch.cuda.synchronize()
    assert (cost != -1).all() and (cost >= 0).all()
    assert cost.max() > 64 and cost.min() == 0   # corner tiles see nothing of the object
