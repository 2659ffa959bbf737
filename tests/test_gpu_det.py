"""Deterministic (sorted-segment) gradient scatter of the fused adjoint:
parity with the float64 oracle, agreement with the atomic modes, and
bitwise equality of two runs, on full frames, on random ray batches across
poses (the training-step form) and on listed (proposal-style) samples."""
import numpy as np
import pytest
import torch

from oracle import raygen
from oracle import render as orc
from qsdvr_amd import _lib, fused, proposal, scenes
from tests.helpers import assert_close

pytestmark = pytest.mark.gpu

FWD, BWD = 1e-5, 1e-4


@pytest.fixture(scope="module")
def small():
    w = scenes.WORKLOADS["small"]
    cam = scenes.orbit_cameras(1, w.width, w.height)[0]
    grid, data, gen = scenes.build_grids(w, seed=0)
    target = torch.rand((w.width * w.height, 3), generator=gen)
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    col, dep = orc.render(g64, d64, smp, w.R, w.s, with_depth=True,
                          cam_center=cam.translation.astype(np.float64))
    (orc.mse_loss(col, target.double()) + 0.01 * dep.mean()).backward()
    return w, cam, grid, data, target, (g64.grad, d64.grad)


def _grads(G0, D0, run):
    G = G0.clone().requires_grad_(True)
    D = D0.clone().requires_grad_(True)
    loss = run(G, D)
    loss.backward()
    torch.cuda.synchronize()
    return G.grad.detach().clone(), D.grad.detach().clone()


def test_det_matches_oracle(small):
    w, cam, grid, data, target, (gq, gd) = small
    tgt = target.cuda()

    def run(G, D):
        c, d = fused.render(G, D, [cam], w.interval, w.s, scatter_mode=fused.SCATTER_DET)
        return ((c - tgt) ** 2).mean() + 0.01 * d.mean()

    a = _grads(grid.cuda(), data.cuda(), run)
    assert_close(a[0].reshape(-1, 7), gq.reshape(-1, 7), BWD, "det sdf-grid grad")
    assert_close(a[1], gd, BWD, "det appearance grad")
    b = _grads(grid.cuda(), data.cuda(), run)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), "two runs differ"


@pytest.fixture(scope="module")
def batch128():
    """65,536 random rays over 20 poses of the 128^3 torus (800x800, N=192)."""
    w = scenes.WORKLOADS["plain"]
    grid, data, gen = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(20, w.width, w.height)
    cam_buf = fused.camera_buffer(cams)
    g = torch.Generator(device="cuda").manual_seed(3)
    n_all = len(cams) * w.width * w.height
    pick = torch.randint(0, n_all, (65536,), generator=g, device="cuda").to(torch.int32)
    target = torch.rand((65536, 3), generator=torch.Generator().manual_seed(4)).cuda()
    return w, grid, data, cams, cam_buf, pick, target


def test_det_random_batch_matches_atomic_and_repeats(batch128):
    w, grid, data, cams, cam_buf, pick, target = batch128

    def runner(mode):
        def run(G, D):
            c, _ = fused.FusedRender.apply(G, D, cam_buf, len(cams), w.width, w.height, pick,
                                           w.interval, w.s, mode)
            return ((c - target) ** 2).mean()
        return run

    det1 = _grads(grid, data, runner(fused.SCATTER_DET))
    det2 = _grads(grid, data, runner(fused.SCATTER_DET))
    assert torch.equal(det1[0], det2[0]), "sdf-grid gradient differs between two runs"
    assert torch.equal(det1[1], det2[1]), "appearance gradient differs between two runs"
    ato = _grads(grid, data, runner(fused.SCATTER_ATOMIC))
    assert_close(det1[0].reshape(-1, 7), ato[0].reshape(-1, 7), BWD, "det vs atomic sdf grad")
    assert_close(det1[1], ato[1], BWD, "det vs atomic appearance grad")
    assert float(det1[1].abs().sum()) > 0 and float(det1[0].abs().sum()) > 0


def test_det_full_frame_matches_tile(batch128):
    w, grid, data, cams, cam_buf, pick, target = batch128
    tgt = torch.rand((w.width * w.height, 3), generator=torch.Generator().manual_seed(5)).cuda()

    def runner(mode):
        def run(G, D):
            c, _ = fused.FusedRender.apply(G, D, cam_buf[80 * 7:80 * 8], 1, w.width, w.height,
                                           None, w.interval, w.s, mode)
            return ((c - tgt) ** 2).mean()
        return run

    det = _grads(grid, data, runner(fused.SCATTER_DET))
    tile = _grads(grid, data, runner(fused.SCATTER_TILE))
    assert_close(det[0].reshape(-1, 7), tile[0].reshape(-1, 7), BWD, "det vs tile sdf grad")
    assert_close(det[1], tile[1], BWD, "det vs tile appearance grad")
    det2 = _grads(grid, data, runner(fused.SCATTER_DET))
    assert torch.equal(det[0], det2[0]) and torch.equal(det[1], det2[1])


def test_det_listed_samples(batch128):
    w, grid, data, cams, cam_buf, pick, target = batch128
    sub = pick[:16384]
    rs = proposal.ray_state(cam_buf, len(cams), w.width, w.height, w.interval, sub)
    cnt = proposal.ray_counts(rs).to(torch.int64)
    L = (torch.clamp(cnt - 1, min=0).float() * w.interval)
    K = 24
    u = torch.sort(torch.rand((sub.shape[0], K - 2), generator=torch.Generator(device="cuda")
                              .manual_seed(6), device="cuda"), dim=1).values
    t = torch.cat([torch.zeros_like(L)[:, None], u * L[:, None], L[:, None]], 1).contiguous()

    def runner(mode):
        def run(G, D):
            c, _ = proposal.ListedRender.apply(G, D, cam_buf, len(cams), w.width, w.height, sub,
                                               rs, t, w.interval, w.s, mode)
            return ((c - target[:16384]) ** 2).mean()
        return run

    det = _grads(grid, data, runner(fused.SCATTER_DET))
    ato = _grads(grid, data, runner(fused.SCATTER_ATOMIC))
    assert_close(det[0].reshape(-1, 7), ato[0].reshape(-1, 7), BWD, "listed det vs atomic sdf")
    assert_close(det[1], ato[1], BWD, "listed det vs atomic appearance")
    det2 = _grads(grid, data, runner(fused.SCATTER_DET))
    assert torch.equal(det[0], det2[0]) and torch.equal(det[1], det2[1])


def test_det_rejects_a_short_workspace(batch128):
    w, grid, data, cams, cam_buf, pick, target = batch128
    sub = pick[:4096]
    rs = proposal.ray_state(cam_buf, len(cams), w.width, w.height, w.interval, sub)
    color = torch.zeros((4096, 3), device="cuda")
    gq = torch.zeros_like(grid)
    gd = torch.zeros_like(data)
    cap = 10  # far fewer slots than the rays' samples
    nb = int(_lib.fn("qsdvr_fused_det_workspace_size")(4096, cap, 32, w.R))
    ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
    from qsdvr_amd.qsdvr import _ptr, _stream
    with pytest.raises(_lib.QsdvrError, match="EWORKSPACE"):
        _lib.call("qsdvr_fused_backward_det", _ptr(cam_buf), len(cams), w.width, w.height,
                  _ptr(sub), 4096, _ptr(rs), None, 0, _ptr(grid), _ptr(data), w.R, 32,
                  float(w.interval), float(w.s), _ptr(color), None, _ptr(color), None, _ptr(gq),
                  _ptr(gd), cap, _ptr(ws), nb, _stream())
    assert float(gq.abs().sum()) == 0.0 and float(gd.abs().sum()) == 0.0
