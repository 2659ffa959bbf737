"""Proposal path on the MI355X against the float64 oracle (oracle/mlp.py):
MFMA forward logits, hand-written backward, sampler, listed-sample fused
render and its adjoint, the bin targets, and the gain of the trained
network over uniform sampling at the same sample count.

Tolerances: logits and parameter gradients are judged scale-aware at 3e-3
(bf16 activations: a pre-activation that lands within float32 rounding of a
bf16 rounding boundary may round the other way); render colours 1e-5,
grid gradients 1e-4 as for the plain path; sample positions 2e-5 of the
segment length (float32 CDF)."""
import numpy as np
import pytest
import torch

from oracle import mlp as om
from oracle import render as orc
from qsdvr_amd import fused, proposal, scenes
from tests.helpers import assert_close
from tests.subset_oracle import listed_lists

pytestmark = pytest.mark.gpu


def _cams(n=4, W=48, H=40):
    return scenes.orbit_cameras(n, W, H)


def _state(cams, interval):
    buf = fused.camera_buffer(cams)
    W, H = cams[0].width, cams[0].height
    rs = proposal.ray_state(buf, len(cams), W, H, interval)
    return buf, rs


def _oracle_points(rs, interval):
    r = rs.double().cpu()
    cnt = proposal.ray_counts(rs).cpu().long()
    return om.bin_centres(r[:, 0:3], r[:, 4:7], cnt, interval).reshape(-1, 3), cnt


def test_mlp_forward_matches_oracle():
    interval = scenes.interval_for(96)
    cams = _cams()
    _, rs = _state(cams, interval)
    prop = proposal.Proposal(seed=3)
    z = prop.logits(rs, interval)
    pts, cnt = _oracle_points(rs, interval)
    ref = om.forward(prop.theta.double().cpu(), pts).reshape(-1, 64)
    live = cnt >= 2
    assert_close(z.cpu()[live].reshape(-1), ref[live].reshape(-1), 3e-3, "proposal logits")


def test_mlp_backward_matches_oracle():
    interval = scenes.interval_for(96)
    cams = _cams(2, 32, 24)
    _, rs = _state(cams, interval)
    n = rs.shape[0]
    g = torch.Generator().manual_seed(5)
    q = torch.rand((n, 64), generator=g) ** 4
    q[::7] = 0.0
    prop = proposal.Proposal(seed=4)
    loss, grad, z = prop.grad(rs, interval, q.cuda())
    pts, cnt = _oracle_points(rs, interval)
    th = prop.theta.double().cpu().requires_grad_(True)
    zr = om.forward(th, pts).reshape(-1, 64)
    lref = om.ray_loss(zr, q.double())
    lref.backward()
    assert abs(float(loss.sum()) - float(lref)) <= 3e-3 * abs(float(lref))
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in om.SHAPES])
    # the output layer is judged as one block [w4, b4]: the b4 gradient,
    # sum over points of Q softmax - q, is exactly zero in exact arithmetic
    blocks = [(n, a, b) for (n, _), a, b in zip(om.SHAPES, offs[:-1], offs[1:]) if n != "b4"]
    blocks[-1] = ("w4+b4", blocks[-1][1], int(offs[-1]))
    for name, a, b in blocks:
        assert_close(grad.cpu()[a:b], th.grad[a:b], 3e-3, f"proposal grad {name}")


def test_sampler_matches_oracle():
    interval = scenes.interval_for(96)
    cams = _cams(2, 30, 20)
    _, rs = _state(cams, interval)
    n = rs.shape[0]
    z = torch.randn((n, 64), generator=torch.Generator().manual_seed(6)) * 3
    cnt = proposal.ray_counts(rs).cpu().long()
    L = torch.clamp(cnt - 1, min=0).double() * np.float32(interval)
    for eps, gamma in ((0.05, 1.0), (0.0, 2.0), (0.1, 1.5)):
        t = proposal.sample(z.cuda(), rs, interval, K=32, eps=eps, gamma=gamma).cpu().double()
        ref = om.fine_samples(z, L, K=32, eps=eps, gamma=gamma)
        err = ((t - ref).abs().max(1).values / L.clamp_min(1e-9))
        assert float(err[cnt >= 2].max()) < 2e-5, (eps, gamma)
        assert torch.all(t[cnt >= 2][:, 1:] >= t[cnt >= 2][:, :-1])


_listed_lists = listed_lists


def _listed_case(mode, sh_degree, K):
    w = scenes.Workload("t", 24, sh_degree, 32, 24, 64, 3, "torus", 0.005)
    grid, data, gen = scenes.build_grids(w, seed=1)
    cams = scenes.orbit_cameras(3, w.width, w.height)[1:2]
    buf, rs = _state(cams, w.interval)
    n = rs.shape[0]
    z = torch.randn((n, 64), generator=torch.Generator().manual_seed(7))
    if K >= 3:
        t = proposal.sample(z.cuda(), rs, w.interval, K=K)
    else:  # the sampler's lists start at 3 samples; the renderer takes 2 (the endpoints)
        t = proposal.sample(z.cuda(), rs, w.interval, K=3)[:, [0, 2]].contiguous()
    s = 8.0
    G = grid.cuda().requires_grad_(True)
    D = data.cuda().requires_grad_(True)
    c, d = proposal.ListedRender.apply(G, D, buf, 1, w.width, w.height, None, rs, t, w.interval,
                                       s, mode)
    target = torch.rand((n, 3), generator=gen)
    (((c - target.cuda()) ** 2).mean() + 0.01 * d.mean()).backward()
    smp = _listed_lists(rs, t, w.R)
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    ref, dep = orc.render(g64, d64, smp, w.R, s, with_depth=True,
                          cam_center=cams[0].translation.astype(np.float64))
    (orc.mse_loss(ref, target.double()) + 0.01 * dep.mean()).backward()
    assert_close(c, ref, 1e-5, f"listed colour (K={K})")
    assert_close(d, dep, 1e-5, f"listed depth (K={K})")
    assert_close(G.grad.reshape(-1, 7), g64.grad.reshape(-1, 7), 1e-4, f"listed sdf grad (K={K})")
    assert_close(D.grad, d64.grad, 1e-4, f"listed appearance grad (K={K})")


@pytest.mark.parametrize("mode", [fused.SCATTER_TILE, fused.SCATTER_ATOMIC, fused.SCATTER_DET])
@pytest.mark.parametrize("sh_degree", [2, 3])
def test_listed_render_matches_oracle(mode, sh_degree):
    """Colour, depth and both grids' gradients of the listed-sample renderer
    (32 or 53 appearance channels) against the float64 oracle on the same
    sample lists."""
    _listed_case(mode, sh_degree, 20)


@pytest.mark.parametrize("mode", [fused.SCATTER_TILE, fused.SCATTER_ATOMIC])
@pytest.mark.parametrize("K", [2, 3])
def test_listed_shortest_lists(mode, K):
    """The shortest lists (one or two render samples per ray), where the
    sample positions and quadrics held a step ahead run out first."""
    _listed_case(mode, 2, K)


@pytest.mark.parametrize("sh_degree", [2, 3])
def test_listed_cost_ordered_matches_launch_order(monkeypatch, sh_degree):
    """Whole-frame listed renders in tile mode run their adjoint heaviest
    tile first (the costed forward and the ordered backward): the same
    colours and depths bit for bit as the launch order, and gradients to
    1e-4 (float atomics in another order), on two poses in one call, at 32
    and 53 appearance channels (53 is not ordered by default)."""
    w = scenes.Workload("t", 24, sh_degree, 96, 80, 64, 3, "torus", 0.005)
    grid, data, gen = scenes.build_grids(w, seed=3)
    cams = scenes.orbit_cameras(3, w.width, w.height)[:2]
    buf, rs = _state(cams, w.interval)
    n = rs.shape[0]
    z = torch.randn((n, 64), generator=torch.Generator().manual_seed(9))
    t = proposal.sample(z.cuda(), rs, w.interval, K=24)
    target = torch.rand((n, 3), generator=gen).cuda()
    monkeypatch.setattr(fused, "TILE_ORDER_CHANNELS", (32, 53))
    out = {}
    for order in (False, True):
        monkeypatch.setattr(fused, "TILE_ORDER", order)
        assert fused._ordered(None, fused.SCATTER_TILE, data.shape[-1]) == order
        G = grid.cuda().requires_grad_(True)
        D = data.cuda().requires_grad_(True)
        c, d = proposal.ListedRender.apply(G, D, buf, 2, w.width, w.height, None, rs, t,
                                           w.interval, 8.0, fused.SCATTER_TILE)
        (((c - target) ** 2).mean() + 0.01 * d.mean()).backward()
        out[order] = (c.detach(), d.detach(), G.grad, D.grad)
    assert torch.equal(out[False][0], out[True][0]) and torch.equal(out[False][1], out[True][1])
    assert_close(out[True][2].reshape(-1, 7), out[False][2].reshape(-1, 7), 1e-4,
                 "ordered vs launch-order listed sdf grad")
    assert_close(out[True][3], out[False][3], 1e-4, "ordered vs launch-order listed appearance grad")


def test_listed_uniform_is_plain_and_bins():
    w = scenes.Workload("t", 32, 2, 40, 30, 80, 2, "torus", 0.005)
    grid, data, _ = scenes.build_grids(w, seed=2)
    cams = scenes.orbit_cameras(2, w.width, w.height)
    buf, rs = _state(cams, w.interval)
    G, D = grid.cuda(), data.cuda()
    c_plain, d_plain = fused.render(G, D, cams, w.interval, w.s)
    c_list, d_list = proposal.ListedRender.apply(G, D, buf, 2, w.width, w.height, None, rs, None,
                                                 w.interval, w.s)
    assert torch.equal(c_plain, c_list) and torch.equal(d_plain, d_list)
    c_b, _, bins = proposal.plain_bins(G, D, buf, 2, w.width, w.height, w.interval, w.s)
    assert torch.equal(c_b, c_plain)
    # bins hold every compositing weight: their sum is the ray opacity
    from oracle import raygen
    smp = raygen.ray_samples(cams[0].as_dict(), np.float32(w.interval), w.R)
    col, wts = orc.composite(*_rgb_sdf(grid, data, smp, w), smp["rayList"], w.s,
                             return_weights=True)
    rc = torch.as_tensor(smp["rayList"][:, 2].astype(np.int64))
    q = om.bin_targets(wts, rc)
    P = w.width * w.height
    assert_close(bins[:P].cpu(), q, 1e-4, "bin targets")
    # the SDF-only bins kernel gives the same targets (T multiplied in scan order)
    sdf_bins = proposal.plain_bins_sdf(G, rs, w.interval, w.s)
    assert_close(sdf_bins.cpu(), bins.cpu(), 1e-5, "SDF-only bins vs fused-forward bins")
    err = (sdf_bins - bins).abs().max().item()
    assert err <= 1e-5, err


def _rgb_sdf(grid, data, smp, w):
    g64, d64 = grid.double(), data.double()
    sdf = orc.sample_sdf(g64, torch.as_tensor(smp["sdfPointList"]).double(),
                         torch.as_tensor(smp["sdfIndexList"]), w.R)
    nrm = orc.sample_normal(g64, torch.as_tensor(smp["renderPointList"]).double(),
                            torch.as_tensor(smp["renderIndexList"]))
    dl = orc.grid_interpolation(d64, torch.as_tensor(smp["renderPointList"]).double(),
                                torch.as_tensor(smp["renderIndexList"]), w.R)
    rgb, _ = orc.shade(nrm, torch.as_tensor(smp["viewDirList"]).double(), dl)
    return rgb, sdf


def test_training_gains_over_uniform_sampling():
    """Trained proposal (32 fine samples) against the plain renderer at 192
    samples (at most 0.1 dB worse), the plain uniform renderer at the same 32
    samples and the untrained network with the same sampler, on held-out
    poses."""
    s = 40.0
    w = scenes.Workload("gain", 64, 2, 96, 96, 192, 12, "torus", 0.005, s=s)
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    train_cams, test_cams = cams[:9], cams[9:]
    prop = proposal.Proposal(seed=0)
    untrained = proposal.Proposal(seed=0)
    hist = proposal.train(prop, grid, data, train_cams, w.interval, s, steps=400, batch=4096,
                          log_every=100)
    assert hist[-1][1] < hist[0][1]
    buf = fused.camera_buffer(test_cams)
    W, H, n = w.width, w.height, len(test_cams)
    gt, _ = fused.render(grid, data, test_cams, w.interval / 8, s)
    u32, _ = fused.render(grid, data, test_cams, 2 * np.sqrt(3) / 32, s)
    p192, _ = fused.render(grid, data, test_cams, w.interval, s)
    c_tr, _ = proposal.render_proposal(grid, data, buf, n, W, H, w.interval, s, prop, K=32)
    c_un, _ = proposal.render_proposal(grid, data, buf, n, W, H, w.interval, s, untrained, K=32)
    p_tr, p_u, p_un, p_pl = (proposal.psnr(c, gt) for c in (c_tr, u32, c_un, p192))
    print(f"PSNR vs GT: trained proposal {p_tr:.2f} dB, plain-192 {p_pl:.2f} dB, "
          f"uniform-32 {p_u:.2f} dB, untrained {p_un:.2f} dB")
    assert p_tr >= p_pl - 0.1
    assert p_tr > p_u + 3.0 and p_tr > p_un + 1.0
