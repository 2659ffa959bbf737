"""Writes a small NeRF-synthetic dataset (transforms_<split>.json plus RGBA
PNG frames) from OpenCV cameras and RGBA images, the inverse of the
reference's loader conventions (datasets/nerf_dataset.py:34-73)."""
import json
import math
import os

import numpy as np

from qsdvr_amd import nerf


def gl_matrix(rotation, translation):
    """4x4 OpenGL camera-to-world matrix of an OpenCV camera (rotation
    row-major [9], translation [3]); diag(1,-1,-1,1) is its own inverse."""
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = np.asarray(rotation, dtype=np.float32).reshape(3, 3)
    m[:3, 3] = np.asarray(translation, dtype=np.float32)
    return m @ nerf.GL_TO_CV


def write_dataset(root, cams, images, split="train"):
    """cams: objects with rotation, translation, width, fx; images: float
    [N, H, W, 4] in [0, 1] (alpha 0 = not part of the object)."""
    os.makedirs(os.path.join(root, split), exist_ok=True)
    frames = []
    for i, (c, img) in enumerate(zip(cams, images)):
        name = f"r_{i}"
        nerf.write_image(os.path.join(root, split, name + ".png"), img)
        frames.append({"file_path": f"./{split}/{name}",
                       "transform_matrix": gl_matrix(c.rotation, c.translation).tolist()})
    angle = 2.0 * math.atan(0.5 * cams[0].width / cams[0].fx)
    with open(os.path.join(root, f"transforms_{split}.json"), "w") as f:
        json.dump({"camera_angle_x": angle, "frames": frames}, f)
