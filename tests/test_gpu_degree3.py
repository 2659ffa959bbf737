"""SH degree 3 (16 coefficients x RGB + 5 = 53 appearance channels; no
reference counterpart, the oracle's degree-3 basis is pinned in
tests/test_oracle_render.py) on the MI355X against the float64 oracle: the
Shader and GridInterpolation ops at C = 53 and the fused renderer, both
scatter modes, on a small gyroid-shell workload.

Tolerances as for degree 2 (scale-aware, tests/helpers.py): forward values
1e-5, gradients 1e-4."""
import numpy as np
import pytest
import torch

from oracle import raygen
from oracle import render as orc
from qsdvr_amd import fused, qsdvr, scenes
from tests.helpers import assert_close

pytestmark = pytest.mark.gpu

FWD, BWD = 1e-5, 1e-4
C = 53


def _f64(t):
    return t.detach().double().cpu().requires_grad_(True)


def test_shader_53():
    g = torch.Generator().manual_seed(31)
    n = 40000
    nrm = torch.nn.functional.normalize(torch.randn((n, 3), generator=g), dim=1)
    view = torch.nn.functional.normalize(torch.randn((n, 3), generator=g), dim=1)
    data = torch.randn((n, C), generator=g) * 0.3
    data[:, C - 5:] = torch.rand((n, 5), generator=g)
    out = torch.empty((n, 3), device="cuda")
    spec = torch.empty((n, 3), device="cuda")
    qsdvr.ShaderForward(out, nrm.cuda(), view.cuda(), data.cuda(), spec)
    n64, d64 = _f64(nrm), _f64(data)
    rgb, L = orc.shade(n64, view.double(), d64)
    assert_close(out, rgb, FWD, "Shader rgb (C=53)")
    assert_close(spec, L, FWD, "Shader specular (C=53)")
    up = torch.randn((n, 3), generator=g)
    (rgb * up.double()).sum().backward()
    gN = torch.empty((n, 3), device="cuda")
    gD = torch.empty((n, C), device="cuda")
    qsdvr.ShaderBackward(up.cuda(), nrm.cuda(), view.cuda(), data.cuda(), spec, gN, gD)
    assert_close(gN, n64.grad, BWD, "Shader normal grad (C=53)")
    assert_close(gD, d64.grad, BWD, "Shader data grad (C=53)")


@pytest.fixture(scope="module")
def gyroid():
    w = scenes.Workload("gyroid24_deg3", 24, 3, 40, 32, 64, 1, "gyroid", 0.005)
    cam = scenes.orbit_cameras(1, w.width, w.height)[0]
    grid, data, gen = scenes.build_grids(w, seed=2)
    assert data.shape[1] == C
    target = torch.rand((w.width * w.height, 3), generator=gen)
    smp = raygen.ray_samples(cam.as_dict(), np.float32(w.interval), w.R)
    return w, cam, grid, data, target, smp


def test_grid_interpolation_53(gyroid):
    w, cam, grid, data, target, smp = gyroid
    R = w.R
    pts = torch.as_tensor(smp["renderPointList"])
    idx = torch.as_tensor(smp["renderIndexList"])
    n = pts.shape[0]
    out = torch.empty((n, C), device="cuda")
    qsdvr.GridInterpolationForward(data.cuda(), pts.cuda(), idx.cuda(), out, R + 1)
    d64 = _f64(data)
    ref = orc.grid_interpolation(d64, pts.double(), idx, R)
    assert_close(out, ref, FWD, "GridInterpolation forward (C=53)")
    up = torch.randn((n, C), generator=torch.Generator().manual_seed(32))
    (ref * up.double()).sum().backward()
    gg = torch.zeros_like(data).cuda()
    qsdvr.GridInterpolationBackward(gg, pts.cuda(), idx.cuda(), up.cuda(), R + 1)
    assert_close(gg, d64.grad, BWD, "GridInterpolation backward (C=53)")


@pytest.mark.parametrize("mode", [fused.SCATTER_TILE, fused.SCATTER_ATOMIC])
def test_fused_53(gyroid, mode):
    w, cam, grid, data, target, smp = gyroid
    g64 = grid.double().requires_grad_(True)
    d64 = data.double().requires_grad_(True)
    col, dep = orc.render(g64, d64, smp, w.R, w.s, with_depth=True,
                          cam_center=cam.translation.astype(np.float64))
    (orc.mse_loss(col, target.double()) + 0.01 * dep.mean()).backward()
    G = grid.cuda().requires_grad_(True)
    D = data.cuda().requires_grad_(True)
    c, d = fused.render(G, D, [cam], w.interval, w.s, scatter_mode=mode)
    assert_close(c, col.detach(), FWD, "fused colour (C=53)")
    assert_close(d, dep.detach(), FWD, "fused depth (C=53)")
    (((c - target.cuda()) ** 2).mean() + 0.01 * d.mean()).backward()
    assert_close(G.grad.reshape(-1, 7), g64.grad.reshape(-1, 7), BWD, "fused sdf-grid grad (C=53)")
    assert_close(D.grad, d64.grad, BWD, "fused appearance grad (C=53)")
