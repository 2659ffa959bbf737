"""NeRF-synthetic format on the host (qsdvr_amd.nerf.load_frames): the
reference loader's conventions, checked against what they must mean
geometrically (datasets/nerf_dataset.py:34-73)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import raygen
from qsdvr_amd import nerf


def _frame_dir(tmp_path, mats, W=8, H=6, angle=0.69):
    os.makedirs(tmp_path / "train")
    rng = np.random.default_rng(0)
    imgs = []
    for i in range(len(mats)):
        img = rng.integers(0, 256, (H, W, 4)).astype(np.uint8)
        img[: H // 2, : W // 2, 3] = 0  # a transparent corner
        nerf._pil().fromarray(img).save(tmp_path / "train" / f"r_{i}.png")
        imgs.append(img)
    frames = [{"file_path": f"./train/r_{i}", "transform_matrix": m} for i, m in enumerate(mats)]
    with open(tmp_path / "transforms_train.json", "w") as f:
        json.dump({"camera_angle_x": angle, "frames": frames}, f)
    return imgs


def test_png_frames_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.random((5, 7, 4))
    nerf.write_image(str(tmp_path / "x.png"), a)
    b = nerf.read_image(str(tmp_path / "x.png"))
    assert b.dtype == np.uint8 and b.shape == (5, 7, 4)
    assert np.array_equal(b, (a * 255).clip(0, 255).astype(np.uint8))
    nerf.write_image(str(tmp_path / "y.png"), a[..., :3])  # RGB: read back opaque
    c = nerf.read_image(str(tmp_path / "y.png"))
    assert np.all(c[..., 3] == 255) and np.array_equal(c[..., :3], b[..., :3])


def test_load_frames_conventions(tmp_path):
    # an OpenGL camera at +4 z looking down -z, and one at +4 x looking down -x
    m0 = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4], [0, 0, 0, 1]]
    m1 = [[0, 0, 1, 4], [0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1]]
    imgs = _frame_dir(tmp_path, [m0, m1])
    fr = nerf.load_frames(str(tmp_path), scene_scale=(0.5, 0.5, 0.5), scene_offset=(0.2, 0, 0))
    assert fr.width == 8 and fr.height == 6 and fr.names == ["r_0", "r_1"]
    assert math.isclose(fr.focal, 0.5 * 8 / math.tan(0.345), rel_tol=1e-12)
    # OpenCV axes: y and z of the camera flipped, centres offset then scaled
    assert np.allclose(fr.c2w[0][:3, :3], np.diag([1, -1, -1]))
    assert np.allclose(fr.c2w[0][:3, 3], [0.1, 0, 2.0])
    assert np.allclose(fr.c2w[1][:3, 3], [2.1, 0, 0])
    assert np.array_equal(fr.images[0], imgs[0]) and np.array_equal(fr.images[1], imgs[1])
    # the flip makes each camera look at the origin: the centre pixel's ray
    # (RayCut restated, CPU) points from the camera centre to the origin
    for i in range(2):
        a = nerf.camera_args(fr.c2w[i], fr.width, fr.height, fr.focal, 0.5, 6.0)
        cam = {"width": a.width, "height": a.height, "rotation": a.rotation,
               "translation": a.translation, "fx": a.fx, "fy": a.fy, "cx": a.cx, "cy": a.cy,
               "near": a.near, "far": a.far}
        k = raygen.ray_cut_kernel(cam, np.float32(0.05))
        d = k["dir"].reshape(fr.height, fr.width, 3)[fr.height // 2, fr.width // 2]
        to_origin = -a.translation / np.linalg.norm(a.translation)
        assert float(np.dot(d, to_origin)) > 0.99, (i, d, to_origin)
    # alpha != 0 marks the rays that are kept; the transparent corner is out
    mask = (fr.images[0][..., 3] != 0)
    assert not mask[:3, :4].any()
    assert np.array_equal(mask, imgs[0][..., 3] != 0)


def test_load_frames_rejects_bad_input(tmp_path):
    os.makedirs(tmp_path / "train")
    with open(tmp_path / "transforms_train.json", "w") as f:
        json.dump({"camera_angle_x": 0.5, "frames": []}, f)
    with pytest.raises(ValueError):
        nerf.load_frames(str(tmp_path))
