"""Oracle parity at the full size of every benchmark workload.

The GPU renders the whole workload (every pixel of a frame, or the whole
65,536-ray batch) with the bench's grids; the float64 oracle then checks a
fixed strided subset of 4,096 of those rays: colour and depth at 1e-5, and
the gradients of a loss over that subset at 1e-4 on every grid row the
subset touches, with both gradient grids exactly zero elsewhere.

* 128^3 torus, SH 2, one full 800x800 frame at N = 192 (tile scatter);
* 256^3 union of 8 spheres, SH 2, a random 65,536-ray batch over 100 poses
  at N = 256 (deterministic, atomic and tile scatter);
* the proposal at 128^3 / 800x800: a trained network's K = 32 fine samples
  per ray through the listed-sample renderer;
* 384^3 gyroid shell, SH 3 (53 channels), one full 1920x1080 frame at
  N = 384: the appearance grid has (R+1)^3 * 53 = 3.0e9 entries, so rows
  past 2^31 / 53 are addressed only with 64-bit offsets, and the subset is
  required to touch some of them.
"""
import numpy as np
import pytest
import torch

from oracle import raygen
from qsdvr_amd import fused, proposal, scenes
from tests.subset_oracle import (camera_samples, check_subset, gpu_subset_loss, listed_lists,
                                 oracle_subset, strided)

pytestmark = pytest.mark.gpu

N_SUB = 4096


def _frame_case(wname, pose, s, seed_target):
    w = scenes.WORKLOADS[wname]
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cam = scenes.orbit_cameras(w.n_poses, w.width, w.height)[pose]
    rows = strided(w.width * w.height, N_SUB)
    target = torch.rand((N_SUB, 3), generator=torch.Generator().manual_seed(seed_target))
    G = grid.requires_grad_(True)
    D = data.requires_grad_(True)
    c, d = fused.render(G, D, [cam], w.interval, s)
    gpu_subset_loss(c, d, rows, target).backward()
    torch.cuda.synchronize()
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R, pixels=rows)
    ref = oracle_subset(G, D, smp, w.R, s, cam.translation.astype(np.float64), target)
    r = torch.as_tensor(rows, device="cuda")
    return w, ref, c[r].detach(), d[r].detach(), G.grad, D.grad


def test_torus128_full_frame():
    w, ref, c, d, gq, gd = _frame_case("plain", 37, scenes.WORKLOADS["plain"].s, 21)
    hit = float((ref["colour"].abs().sum(1) > 0).double().mean())
    print(f"torus128: {len(ref['cells'])} cells, {len(ref['verts'])} vertex rows, "
          f"{hit:.2f} of the rays hit", flush=True)
    assert hit > 0.2
    check_subset(ref, c, d, gq, gd, "torus128 800x800")


def test_gyroid384_sh3_full_frame():
    w, ref, c, d, gq, gd = _frame_case("large", 47, 40.0, 22)
    C = gd.shape[-1]
    assert C == 53
    far = int((torch.as_tensor(ref["verts"]) * C >= 2 ** 31).sum())
    print(f"gyroid384: {len(ref['cells'])} cells, {len(ref['verts'])} vertex rows, "
          f"{far} of them past 2^31 channel entries", flush=True)
    assert far > 1000
    check_subset(ref, c, d, gq, gd, "gyroid384 1920x1080 SH3")


@pytest.fixture(scope="module")
def spheres256():
    w = scenes.WORKLOADS["train"]
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    B = 65536
    n_all = w.n_poses * w.width * w.height
    g = torch.Generator().manual_seed(7)
    codes = torch.randint(0, n_all, (B,), generator=g).to(torch.int32)
    pos = strided(B, N_SUB)
    target = torch.rand((N_SUB, 3), generator=g)
    smp, order, centres = camera_samples(cams, codes.numpy()[pos].astype(np.int64), w.width,
                                         w.height, w.interval, w.R)
    ref = oracle_subset(grid, data, smp, w.R, w.s, centres, target[torch.as_tensor(order)])
    return w, grid, data, fused.camera_buffer(cams), codes.cuda(), pos, order, target, ref


@pytest.mark.parametrize("mode", [fused.SCATTER_DET, fused.SCATTER_ATOMIC, fused.SCATTER_TILE])
def test_spheres256_random_batch(spheres256, mode):
    w, grid, data, cam_buf, codes, pos, order, target, ref = spheres256
    G = grid.detach().requires_grad_(True)
    D = data.detach().requires_grad_(True)
    c, d = fused.FusedRender.apply(G, D, cam_buf, w.n_poses, w.width, w.height, codes,
                                   w.interval, w.s, mode)
    gpu_subset_loss(c, d, pos, target).backward()
    torch.cuda.synchronize()
    r = torch.as_tensor(pos[order], device="cuda")
    check_subset(ref, c[r].detach(), d[r].detach(), G.grad, D.grad,
                 f"spheres256 65,536 rays (scatter {mode})")


def test_proposal_k32_full_frame():
    w = scenes.WORKLOADS["plain"]
    s = 40.0
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    prop = proposal.Proposal(seed=0)
    proposal.train(prop, grid, data, cams[:90], w.interval, s, steps=300, batch=8192)
    cam = cams[93]
    cb = fused.camera_buffer([cam])
    W, H = w.width, w.height
    rs = proposal.ray_state(cb, 1, W, H, w.interval)
    t = proposal.sample(prop.logits(rs, w.interval), rs, w.interval, K=32)
    G = grid.requires_grad_(True)
    D = data.requires_grad_(True)
    c, d = proposal.ListedRender.apply(G, D, cb, 1, W, H, None, rs, t, w.interval, s,
                                       fused.SCATTER_TILE)
    rows = strided(W * H, N_SUB)
    target = torch.rand((N_SUB, 3), generator=torch.Generator().manual_seed(23))
    gpu_subset_loss(c, d, rows, target).backward()
    torch.cuda.synchronize()
    r = torch.as_tensor(rows, device="cuda")
    smp = listed_lists(rs[r], t[r], w.R)
    ref = oracle_subset(G, D, smp, w.R, s, cam.translation.astype(np.float64), target)
    check_subset(ref, c[r].detach(), d[r].detach(), G.grad, D.grad, "proposal K=32 800x800")
