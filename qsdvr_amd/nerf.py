"""NeRF-synthetic data on disk -> the render path, and images back out.

The reference's only workload is the NeRF-synthetic format
(datasets/nerf_dataset.py:14-96): ``transforms_<split>.json`` with
``camera_angle_x`` and per-frame 4x4 camera-to-world matrices in OpenGL
axes, and RGBA PNG frames whose alpha marks the object.  ``load_frames``
reads it on the host: OpenGL -> OpenCV by right-multiplying diag(1,-1,-1,1)
(nerf_dataset.py:43-49), focal 0.5 W / tan(camera_angle_x / 2) (:54-55),
the scene offset added to and the scale applied to the camera centres
(:57-58), principal point (-W/2, -H/2) (:63-64), pixels with alpha != 0
kept (:71-73).  ``NerfDataset`` builds the per-frame masked ray lists through
``camera.Camera`` (RayCut with the alpha mask, rays in masked-pixel order,
datasets/utils.py:5-56) and serves (ray-point lists, RGB targets) as the
reference does; ``fused_rays`` gives the same rays, in the same order, to
the fused renderer.  ``get_whole_image`` (nerf_dataset.py:92-96),
``write_image`` (test.py:44-45) and ``psnr`` turn rendered rays back into
images and scores.

Images are read and written with PIL.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import fused, qsdvr
from .camera import Camera

# OpenGL -> OpenCV camera axes (nerf_dataset.py:43)
GL_TO_CV = np.diag(np.array([1.0, -1.0, -1.0, 1.0], dtype=np.float32))


def _pil():
    try:
        from PIL import Image
    except ImportError as e:  # pragma: no cover - PIL ships with this image
        raise RuntimeError("reading and writing PNG frames needs PIL") from e
    return Image


def read_image(path: str) -> np.ndarray:
    """uint8 [H, W, 4] (RGBA; a frame without alpha is fully opaque)."""
    img = np.asarray(_pil().open(path))
    if img.dtype != np.uint8:
        raise ValueError(f"{path}: expected 8-bit channels, got {img.dtype}")
    if img.ndim == 2:
        img = np.repeat(img[..., None], 3, axis=2)
    if img.shape[2] == 3:
        img = np.concatenate([img, np.full(img.shape[:2] + (1,), 255, np.uint8)], axis=2)
    if img.shape[2] != 4:
        raise ValueError(f"{path}: unsupported channel count {img.shape[2]}")
    return img


def write_image(path: str, img) -> None:
    """Writes a float image in [0, 1] (H, W, 3 or 4) as an 8-bit PNG, clipped
    as test.py:44-45 does ((img * 255).clip(0, 255))."""
    a = np.asarray(img, dtype=np.float64)
    _pil().fromarray((a * 255).clip(0, 255).astype(np.uint8)).save(path)


@dataclass
class Frames:
    """One split on the host: OpenCV camera-to-world matrices and images."""
    c2w: np.ndarray        # float32 [N, 4, 4]
    images: np.ndarray     # uint8 [N, H, W, 4]
    focal: float
    names: list

    @property
    def width(self) -> int:
        return int(self.images.shape[2])

    @property
    def height(self) -> int:
        return int(self.images.shape[1])


def load_frames(root: str, split: str = "train", scene_scale=(1.0, 1.0, 1.0),
                scene_offset=(0.0, 0.0, 0.0)) -> Frames:
    """Reads ``<root>/transforms_<split>.json`` and the frames it names
    (``<root>/<split>/<basename(file_path)>.png``, nerf_dataset.py:34-52)."""
    with open(os.path.join(root, f"transforms_{split}.json")) as f:
        meta = json.load(f)
    frames = meta.get("frames", [])
    if not frames:
        raise ValueError(f"{root}: transforms_{split}.json lists no frames")
    c2w, imgs, names = [], [], []
    for fr in frames:
        name = os.path.basename(fr["file_path"])
        m = np.asarray(fr["transform_matrix"], dtype=np.float32).reshape(4, 4)
        c2w.append(m @ GL_TO_CV)
        imgs.append(read_image(os.path.join(root, split, name + ".png")))
        names.append(name)
    shapes = {im.shape for im in imgs}
    if len(shapes) != 1:
        raise ValueError(f"{root}: frames of different sizes {sorted(shapes)}")
    c2w = np.stack(c2w)
    c2w[:, :3, 3] += np.asarray(scene_offset, dtype=np.float32)
    c2w[:, :3, 3] *= np.asarray(scene_scale, dtype=np.float32)
    images = np.stack(imgs)
    focal = float(0.5 * images.shape[2] / np.tan(0.5 * float(meta["camera_angle_x"])))
    return Frames(c2w=c2w, images=images, focal=focal, names=names)


def camera_args(c2w: np.ndarray, width: int, height: int, focal: float, near: float,
                far: float) -> qsdvr.CameraArgs:
    """CameraArgs of one frame (nerf_dataset.py:78-90)."""
    a = qsdvr.CameraArgs()
    a.width = int(width)
    a.height = int(height)
    a.rotation = np.asarray(c2w, dtype=np.float32)[:3, :3].reshape(-1)
    a.translation = np.asarray(c2w, dtype=np.float32)[:3, 3].reshape(-1)
    a.fx = float(focal)
    a.fy = float(focal)
    a.cx = -width / 2
    a.cy = -height / 2
    a.near = float(near)
    a.far = float(far)
    return a


class NerfDataset:
    """datasets/nerf_dataset.py's NerfDataset on this build's ops: item i is
    (ray-point lists of frame i's masked rays, their RGB targets [n, 3] on
    the device).  The per-frame RayCut runs once at construction."""

    def __init__(self, root: str, near: float = 0.5, far: float = 2.5,
                 scene_scale=(1.0, 1.0, 1.0), scene_offset=(0.0, 0.0, 0.0),
                 interval: float = 0.03, reso: int = 128, split: str = "train",
                 device: str = "cuda"):
        self.interval = float(interval)
        self.reso = int(reso)
        self.near, self.far = float(near), float(far)
        self.frames = load_frames(root, split, scene_scale, scene_offset)
        self.width, self.height = self.frames.width, self.frames.height
        self.n_images = len(self.frames.names)
        self.focal = self.frames.focal
        self.c2w = self.frames.c2w
        self.cx, self.cy = -self.width / 2, -self.height / 2
        self.device = device
        self._cam_buf = None
        self.gt, self.mask_index, self.cameras = [], [], []
        for i in range(self.n_images):
            img = torch.from_numpy(self.frames.images[i].reshape(-1, 4).astype(np.float32) / 255.0)
            mask = img[:, 3] != 0
            self.mask_index.append(mask.nonzero().reshape(-1).numpy())
            self.gt.append(img[mask, :3].to(device))
            cam = self.create_camera(self.c2w[i])
            cam.GenerateRayPoints(mask)
            self.cameras.append(cam)

    def create_camera(self, c2w) -> Camera:
        args = camera_args(c2w, self.width, self.height, self.focal, self.near, self.far)
        return Camera(args, self.interval, self.reso)

    def fused_rays(self, index: int):
        """(camera buffer, int32 pixel codes) of frame ``index``'s masked rays
        for ``fused.FusedRender`` (one camera; rows in masked-pixel order,
        the order of ``self.gt[index]``)."""
        buf = fused.camera_buffer([self.cameras[index].args], self.device)
        codes = torch.as_tensor(self.mask_index[index].astype(np.int32), device=self.device)
        return buf, codes

    def camera_buffer(self) -> torch.Tensor:
        """Device camera buffer of every frame (row i = frame i), built once:
        the ``cam_buf`` of a ``train.FusedTrainer`` over this dataset."""
        if self._cam_buf is None:
            if self.n_images * self.width * self.height >= 2 ** 31:
                raise ValueError("ray codes of this dataset do not fit in int32")
            self._cam_buf = fused.camera_buffer([c.args for c in self.cameras], self.device)
        return self._cam_buf

    def frame_rays(self, index: int) -> torch.Tensor:
        """int32 codes ``index * W * H + pixel`` of frame ``index``'s masked
        rays against ``camera_buffer()``, in the order of ``self.gt[index]``:
        ``trainer.step(ds.frame_rays(i), ds.gt[i])`` is one step of
        test.py:30-35 (all masked rays of one frame per step)."""
        self.camera_buffer()
        base = index * self.width * self.height
        return torch.as_tensor((self.mask_index[index] + base).astype(np.int32),
                               device=self.device)

    def get_whole_image(self, sparse_image, index: int) -> np.ndarray:
        """Scatters the masked rays' colours back into an [H, W, 3] image;
        unmasked pixels are 0 (nerf_dataset.py:92-96)."""
        img = np.zeros((self.width * self.height, 3))
        img[self.mask_index[index]] = np.asarray(sparse_image)
        return img.reshape(self.height, self.width, 3)

    def __len__(self) -> int:
        return self.n_images

    def __getitem__(self, index: int):
        return self.cameras[index].GetRayPoints(), self.gt[index]


def psnr(img, ref, data_range: float = 1.0) -> float:
    """PSNR in dB of an image against a reference in [0, data_range]."""
    a = torch.as_tensor(np.asarray(img), dtype=torch.float64)
    b = torch.as_tensor(np.asarray(ref), dtype=torch.float64)
    mse = float(((a - b) ** 2).mean())
    return math.inf if mse == 0.0 else 10.0 * math.log10(data_range ** 2 / mse)
