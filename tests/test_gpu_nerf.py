"""NeRF-synthetic frames through the render path on the MI355X: a dataset
written in the reference's on-disk format (cameras in OpenGL axes, RGBA
frames whose alpha masks rays) is loaded by qsdvr_amd.nerf.NerfDataset;
the masked rays' op-surface lists and the fused renderer on the same rays
agree (1e-5), and the rendered rays assembled back into images reproduce
the frames up to their 8-bit quantisation."""
import numpy as np
import pytest
import torch

from qsdvr_amd import fused, nerf, scenes
from qsdvr_amd import render as hr
from tests.helpers import assert_close
from tests.nerf_fixture import write_dataset

pytestmark = pytest.mark.gpu


def test_nerf_dataset_through_both_renderers(tmp_path):
    w = scenes.Workload("nerf", 24, 2, 48, 40, 64, 3, "torus", 0.005)
    grid, data, _ = scenes.build_grids(w, seed=3, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    col, _ = fused.render(grid, data, cams, w.interval, w.s)
    rgb = col.clamp(0, 1).reshape(len(cams), w.height, w.width, 3).cpu().numpy()
    alpha = np.ones((len(cams), w.height, w.width, 1))
    alpha[:, ::5, ::3] = 0.0  # masked-out pixels
    write_dataset(str(tmp_path), cams, np.concatenate([rgb, alpha], axis=3))

    ds = nerf.NerfDataset(str(tmp_path), near=cams[0].near, far=cams[0].far,
                          interval=w.interval, reso=w.R)
    assert len(ds) == len(cams) and (ds.width, ds.height) == (w.width, w.height)
    keep = alpha.reshape(len(cams), -1) != 0
    for i in range(len(ds)):
        inp, gt = ds[i]
        assert np.array_equal(ds.mask_index[i], np.nonzero(keep[i])[0])
        assert inp["rayList"].shape[0] == gt.shape[0] == keep[i].sum()
        # op-surface lists of the masked rays vs the fused renderer on them
        out = hr.render_lists(grid, data, inp, w.s)
        buf, codes = ds.fused_rays(i)
        c, _ = fused.FusedRender.apply(grid, data, buf, 1, w.width, w.height, codes,
                                       w.interval, w.s)
        assert_close(out, c, 1e-5, f"frame {i}: op surface vs fused on the masked rays")
        # the same rays as the frame they were rendered into
        assert_close(c, col.reshape(len(cams), -1, 3)[i][torch.as_tensor(keep[i]).cuda()],
                     1e-5, f"frame {i}: masked rays vs the full frame")
        # image assembly: the frame back, up to 8-bit quantisation (the
        # writer truncates, so a channel reads back up to 1/255 low)
        img = ds.get_whole_image(c.clamp(0, 1).cpu().numpy(), i)
        ref = ds.frames.images[i][..., :3].astype(np.float64) / 255.0
        m = ds.frames.images[i][..., 3] != 0
        assert np.all(img[~m] == 0)
        assert nerf.psnr(img[m], ref[m]) > 50.0
        # (the focal length round-trips through camera_angle_x, so these rays
        # may differ from the frame's in the last bits)
        d = img[m] - ref[m]
        assert float(d.min()) >= -1e-5 and float(d.max()) < 1.0 / 255 + 1e-5
        nerf.write_image(str(tmp_path / f"render_{i}.png"), img)
        back = nerf.read_image(str(tmp_path / f"render_{i}.png"))
        codes = np.abs(back[..., :3][m].astype(int) - ds.frames.images[i][..., :3][m])
        assert codes.max() <= 1 and codes.mean() < 0.01


def test_training_over_nerf_frames(tmp_path):
    """test.py:27-35 on the device: per epoch, one Adam step per frame over
    that frame's masked rays, from a perturbed appearance grid towards the
    frames rendered from the unperturbed one."""
    from qsdvr_amd.train import FusedTrainer
    w = scenes.Workload("nerf", 24, 1, 40, 32, 64, 4, "sphere", 0.005)
    grid, data, gen = scenes.build_grids(w, seed=5, device="cuda")
    cams = scenes.orbit_cameras(w.n_poses, w.width, w.height)
    col, _ = fused.render(grid, data, cams, w.interval, w.s)
    rgb = col.clamp(0, 1).reshape(len(cams), w.height, w.width, 3).cpu().numpy()
    alpha = np.ones((len(cams), w.height, w.width, 1))
    alpha[:, 1::4, ::2] = 0.0
    write_dataset(str(tmp_path), cams, np.concatenate([rgb, alpha], axis=3))
    ds = nerf.NerfDataset(str(tmp_path), near=cams[0].near, far=cams[0].far,
                          interval=w.interval, reso=w.R)

    # the all-frames buffer and per-frame codes give the rays the
    # single-frame buffer gives
    for i in range(len(ds)):
        a, _ = fused.FusedRender.apply(grid, data, ds.camera_buffer(), len(ds), ds.width,
                                       ds.height, ds.frame_rays(i), w.interval, w.s)
        buf, codes = ds.fused_rays(i)
        b, _ = fused.FusedRender.apply(grid, data, buf, 1, ds.width, ds.height, codes,
                                       w.interval, w.s)
        assert torch.equal(a, b), f"frame {i}"

    noise = torch.randn(data.shape, generator=gen).to(data.device) * 0.2
    tr = FusedTrainer(data + noise, ds.camera_buffer(), len(ds), ds.width, ds.height,
                      w.interval, w.s, sdf_grid=grid, lr=1e-2, seed=1)
    epochs = []
    for _ in range(6):
        losses = [tr.step(ds.frame_rays(i), ds.gt[i]) for i in range(len(ds))]
        epochs.append(float(torch.stack(losses).mean()))
    assert all(np.isfinite(epochs))
    assert epochs[-1] < 0.5 * epochs[0], epochs
