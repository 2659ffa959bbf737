"""The tile adjoint when a tile's rays diverge: a 150-degree field of view on
a 16x16 frame puts neighbouring rays many cells apart, so a step's box of
touched vertices no longer fits the 64-vertex weight buffer and is taken in
several rounds (slabs of the box, or 3x3x3-cell boxes around single lanes).
Gradients against the float64 oracle and against the per-sample atomic
scatter."""
import numpy as np
import pytest
import torch

from oracle import raygen
from oracle import render as orc
from qsdvr_amd import fused, scenes
from tests.helpers import assert_close

pytestmark = pytest.mark.gpu

FWD, BWD = 1e-5, 1e-4


def _largest_step_box(smp, W, H, R):
    """Largest vertex box of one step of one 8x8 tile over its rays' render
    samples (a bound on the box of the samples with nonzero weight)."""
    rl = smp["rayList"]
    idx = smp["renderIndexList"].astype(np.int64)
    big = 0
    for ty in range(0, H, 8):
        for tx in range(0, W, 8):
            rays = [py * W + px for py in range(ty, min(ty + 8, H)) for px in range(tx, min(tx + 8, W))]
            n = max(int(rl[r, 2]) for r in rays)
            for j in range(n):
                cells = [idx[rl[r, 3] + j] for r in rays if j < rl[r, 2]]
                c = np.array([[q % R, (q // R) % R, q // (R * R)] for q in cells])
                ext = c.max(0) - c.min(0) + 2
                big = max(big, int(ext.prod()))
    return big


@pytest.mark.parametrize("sh_degree", [2, 3])
def test_diverging_tile_rays(sh_degree):
    w = scenes.Workload("wide", 48, sh_degree, 16, 16, 96, 1, "torus", 0.005)
    cam = scenes.orbit_cameras(1, w.width, w.height, radius_world=3.0, fov_deg=150.0)[0]
    grid, data, gen = scenes.build_grids(w, seed=4)
    target = torch.rand((w.width * w.height, 3), generator=gen)
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    assert _largest_step_box(smp, w.width, w.height, w.R) > 64  # rounds are needed
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    col = orc.render(g64, d64, smp, w.R, w.s)
    orc.mse_loss(col, target.double()).backward()
    grads = {}
    for mode in (fused.SCATTER_TILE, fused.SCATTER_ATOMIC):
        G = grid.cuda().requires_grad_(True)
        D = data.cuda().requires_grad_(True)
        c, _ = fused.render(G, D, [cam], w.interval, w.s, scatter_mode=mode)
        assert_close(c, col.detach(), FWD, f"colour (mode {mode})")
        ((c - target.cuda()) ** 2).mean().backward()
        assert_close(G.grad.reshape(-1, 7), g64.grad.reshape(-1, 7), BWD, f"sdf grad (mode {mode})")
        assert_close(D.grad, d64.grad, BWD, f"appearance grad (mode {mode})")
        grads[mode] = (G.grad, D.grad)
    assert_close(grads[fused.SCATTER_TILE][1], grads[fused.SCATTER_ATOMIC][1], BWD,
                 "tile vs atomic appearance grad")
