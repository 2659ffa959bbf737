"""GPU-resident training step: the HIP Adam update against torch.optim.Adam,
loss decrease on a small scene, layer-parameterised grids on the device
against the reference's host layers, and bitwise-repeatable training with
the deterministic scatter."""
import pytest
import torch

from qsdvr_amd import fused, scenes
from qsdvr_amd import render as hr
from qsdvr_amd.train import DeviceAdam, FusedTrainer

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,offset", [(4096, 0), (100003, 0), (4099, 1)])
def test_device_adam_matches_torch_adam(n, offset):
    g = torch.Generator().manual_seed(n)
    base = torch.randn(n + offset, generator=g)
    ref = base[offset:].clone().double().requires_grad_(True)
    buf = base.clone().cuda()
    p = buf[offset:]  # offset 1: a view off 16-byte alignment takes the scalar path
    opt_ref = torch.optim.Adam([ref], lr=1e-2)
    opt = DeviceAdam([p], lr=1e-2)
    for _ in range(3):
        grad = torch.randn(n, generator=g)
        ref.grad = grad.double()
        opt_ref.step()
        p.grad = grad.cuda()
        opt.step()
    torch.cuda.synchronize()
    err = (p.double().cpu() - ref.detach()).abs().max().item()
    assert err <= 1e-6, err


def _small_scene():
    w = scenes.Workload("torus32_64x64_n64", 32, 2, 64, 64, 64, 4, "torus", 0.0)
    grid, data, _ = scenes.build_grids(w, seed=0, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    return w, grid, data, cams, fused.camera_buffer(cams)


def test_trainer_fits_perturbed_appearance():
    w, grid, data, cams, cam_buf = _small_scene()
    n_all = len(cams) * w.width * w.height
    codes = torch.arange(n_all, dtype=torch.int32, device="cuda")
    with torch.no_grad():
        gt, _ = fused.FusedRender.apply(grid, data, cam_buf, len(cams), w.width, w.height, codes,
                                        w.interval, w.s, fused.SCATTER_TILE)
    noisy = data + 0.2 * torch.randn(data.shape, generator=torch.Generator(device="cuda")
                                     .manual_seed(1), device="cuda")
    tr = FusedTrainer(noisy, cam_buf, len(cams), w.width, w.height, w.interval, w.s,
                      sdf_grid=grid.clone(), lr=2e-2)
    losses = []
    for _ in range(40):
        rays = tr.draw(4096)
        losses.append(float(tr.step(rays, gt[rays.long()])))
    assert losses[-1] < 0.3 * losses[0], (losses[0], losses[-1])


def test_device_layers_match_host_layers():
    w, grid, data, cams, cam_buf = _small_scene()
    rays = torch.randint(0, len(cams) * w.width * w.height, (8192,),
                         generator=torch.Generator(device="cuda").manual_seed(2),
                         device="cuda").to(torch.int32)
    target = torch.rand((8192, 3), generator=torch.Generator(device="cuda").manual_seed(3),
                        device="cuda")
    grads = []
    for dev in ("cpu", "cuda"):
        q = hr.QuadricGrid(w.R, device=dev)
        with torch.no_grad():  # a radius-0.7 sphere, slightly tilted layers
            q.init_as_sphere(0.7 * 0.5)
            q.xLayer.mul_(1.0 + 0.01 * torch.arange(w.R, dtype=torch.float32, device=dev) / w.R)
        D = data.clone().requires_grad_(True)
        c, _ = fused.FusedRender.apply(q.grid(), D, cam_buf, len(cams), w.width, w.height, rays,
                                       w.interval, w.s, fused.SCATTER_DET)
        ((c - target) ** 2).mean().backward()
        grads.append([t.grad.detach().cpu() for t in (q.xLayer, q.yLayer, q.zLayer, q.offset)]
                     + [D.grad.detach().cpu()])
    for a, b in zip(*grads):
        assert torch.equal(a, b)
    assert float(grads[0][0].abs().sum()) > 0


def test_deterministic_training_repeats_bitwise():
    w, grid, data, cams, cam_buf = _small_scene()
    target = torch.rand((len(cams) * w.width * w.height, 3),
                        generator=torch.Generator(device="cuda").manual_seed(4), device="cuda")
    finals = []
    for _ in range(2):
        tr = FusedTrainer(data.clone(), cam_buf, len(cams), w.width, w.height, w.interval, w.s,
                          sdf_grid=grid.clone(), lr=1e-2, seed=5)
        for _ in range(3):
            rays = tr.draw(4096)
            tr.step(rays, target[rays.long()])
        torch.cuda.synchronize()
        finals.append((tr.sdf.detach().clone(), tr.data.detach().clone()))
    assert torch.equal(finals[0][0], finals[1][0])
    assert torch.equal(finals[0][1], finals[1][1])


def test_checkpoint_resume_is_bitwise(tmp_path):
    """Three steps, checkpoint, three more steps == resume from the file and
    run the same three steps (deterministic scatter)."""
    w, grid, data, cams, cam_buf = _small_scene()
    target = torch.rand((len(cams) * w.width * w.height, 3),
                        generator=torch.Generator(device="cuda").manual_seed(6), device="cuda")

    def make():
        return FusedTrainer(data.clone(), cam_buf, len(cams), w.width, w.height, w.interval, w.s,
                            sdf_grid=grid.clone(), lr=1e-2, seed=9)

    a = make()
    for _ in range(3):
        r = a.draw(2048)
        a.step(r, target[r.long()])
    path = tmp_path / "trainer.pt"
    torch.save(a.state_dict(), path)
    for _ in range(3):
        r = a.draw(2048)
        a.step(r, target[r.long()])
    b = make()
    b.load_state_dict(torch.load(path, weights_only=True))
    for _ in range(3):
        r = b.draw(2048)
        b.step(r, target[r.long()])
    torch.cuda.synchronize()
    assert torch.equal(a.sdf, b.sdf) and torch.equal(a.data, b.data)
    assert a.opt.t == b.opt.t == 6
