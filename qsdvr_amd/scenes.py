"""Synthetic workloads: cameras, quadric grids and appearance grids.

These generate the *inputs* of the render path (they are not part of the
computation being measured).  Conventions:

* World coordinates [-1.5, 1.5]^3 are mapped onto the reference's hard-coded
  unit cube [-1, 1]^3 (kernel/camera.cu:46, 73) by dividing by 1.5, the role
  ``scene_scale`` plays in datasets/nerf_dataset.py:57-58.  Camera distance,
  SDF values and noise sigma all scale by 1/1.5.
* Cameras follow datasets/nerf_dataset.py:43-90: OpenCV axes (x right,
  y down, z forward), row-major 3x3 camera-to-world rotation, focal
  0.5*W/tan(fov/2), principal point cx=-W/2, cy=-H/2, near/far are
  camera-space depths (here d -+ sqrt(3), bracketing the cube).
* The quadric grid [R,R,R,7] ([z][y][x][A..G], kernel/surface.cu:3-24) is a
  per-cell second-order Taylor fill of the normalised SDF phi at the cell
  centre c, computed in float64: G = R*(phi(c) + sigma*xi),
  (D,E,F) = grad phi(c), (A,B,C) = diag(Hess phi(c)) / (2R).  With the
  local point p in [-1,1), the cube point is c + p/R, so
  f(p)/R = phi(c) + grad phi . p/R + p^2 phi''/(2R^2) + noise.
* The appearance grid [(R+1)^3, C] holds SH coefficients N(0, 0.1) in
  channels 3k+c for the active degree (higher ones zero), diffuse, metallic
  and ao U(0, 1) in channels C-5 .. C-1.
* Random draws use a CPU torch.Generator in the fixed order
  quadric noise -> appearance grid -> loss target.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

WORLD_HALF = 1.5


@dataclass
class Camera:
    """Pinhole camera in the reference's CameraArgs layout
    (include/camera.h:4-34)."""
    width: int
    height: int
    rotation: np.ndarray          # float32 [9], row-major camera-to-world
    translation: np.ndarray       # float32 [3]
    fx: float
    fy: float
    cx: float
    cy: float
    near: float
    far: float

    def as_dict(self) -> dict:
        return {"width": self.width, "height": self.height,
                "rotation": self.rotation, "translation": self.translation,
                "fx": self.fx, "fy": self.fy, "cx": self.cx, "cy": self.cy,
                "near": self.near, "far": self.far}


def fibonacci_directions(n: int) -> np.ndarray:
    """n unit vectors on a Fibonacci lattice (float64 [n,3])."""
    i = np.arange(n, dtype=np.float64) + 0.5
    zc = 1.0 - 2.0 * i / n
    r = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), zc], axis=1)


def look_at_camera(position, width: int, height: int, fov_deg: float = 40.0,
                   near: float | None = None, far: float | None = None) -> Camera:
    """Camera at ``position`` (normalised units) looking at the origin with
    world-z up (nudged at the poles)."""
    pos = np.asarray(position, dtype=np.float64)
    d = float(np.linalg.norm(pos))
    fwd = -pos / d
    up = np.array([0.0, 0.0, 1.0])
    if abs(float(fwd @ up)) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    rot = np.stack([right, down, fwd], axis=1)           # columns: x, y, z axes
    f = 0.5 * width / math.tan(math.radians(fov_deg) / 2)
    s3 = math.sqrt(3.0)
    return Camera(width=width, height=height,
                  rotation=rot.reshape(9).astype(np.float32),
                  translation=pos.astype(np.float32),
                  fx=f, fy=f, cx=-width / 2, cy=-height / 2,
                  near=(d - s3) if near is None else near,
                  far=(d + s3) if far is None else far)


def orbit_cameras(n: int, width: int, height: int, radius_world: float = 4.0,
                  fov_deg: float = 40.0) -> list[Camera]:
    d = radius_world / WORLD_HALF
    return [look_at_camera(d * v, width, height, fov_deg) for v in fibonacci_directions(n)]


def interval_for(n_samples: int) -> float:
    """interval = 2*sqrt(3)/N: the cube diagonal gets N steps."""
    return 2.0 * math.sqrt(3.0) / n_samples


# ---------------------------------------------------------------- SDFs ---
# Each takes float64 [...,3] normalised points and returns phi (normalised).

def sdf_sphere(p, center=(0.0, 0.0, 0.0), radius_world=1.0):
    c = torch.as_tensor(center, dtype=p.dtype, device=p.device)
    return torch.linalg.norm(p - c, dim=-1) - radius_world / WORLD_HALF


def sdf_torus(p, major_world=1.0, minor_world=0.35):
    Rm, rm = major_world / WORLD_HALF, minor_world / WORLD_HALF
    q = torch.sqrt(p[..., 0] ** 2 + p[..., 1] ** 2) - Rm
    return torch.sqrt(q * q + p[..., 2] ** 2) - rm


def make_sphere_union(seed: int = 0, n: int = 8):
    g = np.random.default_rng(seed)
    radii = g.uniform(0.2, 0.6, n) / WORLD_HALF
    centres = g.uniform(-1.0, 1.0, (n, 3)) / WORLD_HALF

    def phi(p):
        out = None
        for c, r in zip(centres, radii):
            d = torch.linalg.norm(p - torch.as_tensor(c, dtype=p.dtype, device=p.device),
                                  dim=-1) - float(r)
            out = d if out is None else torch.minimum(out, d)
        return out
    return phi


def sdf_gyroid_shell(p, thickness=0.2):
    """|sin x cos y + sin y cos z + sin z cos x| - 0.2 with x = 2*pi*w, w the
    world coordinate; the implicit value is divided by its gradient norm so
    it approximates a distance (world units), then normalised."""
    w = p * WORLD_HALF
    x = 2 * math.pi * w
    sx, cx = torch.sin(x[..., 0]), torch.cos(x[..., 0])
    sy, cy = torch.sin(x[..., 1]), torch.cos(x[..., 1])
    sz, cz = torch.sin(x[..., 2]), torch.cos(x[..., 2])
    gval = sx * cy + sy * cz + sz * cx
    gx = 2 * math.pi * (cx * cy - sz * sx)
    gy = 2 * math.pi * (-sx * sy + cy * cz)
    gz = 2 * math.pi * (-sy * sz + cz * cx)
    gn = torch.sqrt(gx * gx + gy * gy + gz * gz).clamp_min(1e-3)
    val = (torch.abs(gval) - thickness) / gn
    return val / WORLD_HALF


SCENES = {"sphere": sdf_sphere, "torus": sdf_torus, "gyroid": sdf_gyroid_shell}


def quadric_grid(phi, R: int, noise_sigma_world: float = 0.0,
                 gen: torch.Generator | None = None, device="cpu",
                 chunk: int = 1 << 21) -> torch.Tensor:
    """Per-cell Taylor fill (float64 maths, float32 result) [R,R,R,7]."""
    h1 = 1e-4
    h2 = 1e-3
    n = R ** 3
    out = torch.empty((n, 7), dtype=torch.float32, device=device)
    sigma = noise_sigma_world / WORLD_HALF
    for s0 in range(0, n, chunk):
        s1 = min(n, s0 + chunk)
        cid = torch.arange(s0, s1, dtype=torch.int64, device=device)
        cx = cid % R
        cy = (cid // R) % R
        cz = cid // (R * R)
        c = (torch.stack([cx, cy, cz], 1).to(torch.float64) * 2 + 1) / R - 1
        f0 = phi(c)
        grads, hess = [], []
        for a in range(3):
            e = torch.zeros(3, dtype=torch.float64, device=device)
            e[a] = h1
            grads.append((phi(c + e) - phi(c - e)) / (2 * h1))
            e[a] = h2
            hess.append((phi(c + e) - 2 * f0 + phi(c - e)) / (h2 * h2))
        if sigma > 0:
            xi = torch.randn(s1 - s0, generator=gen, dtype=torch.float64).to(device)
            f0 = f0 + sigma * xi
        blk = torch.stack([hess[0] / (2 * R), hess[1] / (2 * R), hess[2] / (2 * R),
                           grads[0], grads[1], grads[2], R * f0], dim=1)
        out[s0:s1] = blk.to(torch.float32)
    return out.reshape(R, R, R, 7)


def channels_for_degree(deg: int) -> int:
    """Appearance channel count: (deg+1)^2 SH coefficients x RGB + diffuse
    RGB + metallic + ao.  Degrees 1 and 2 share the reference's 32-channel
    layout (degree 1 leaves coefficients 4..8 at zero)."""
    if deg <= 2:
        return 32
    return 3 * (deg + 1) ** 2 + 5


def appearance_grid(R: int, sh_degree: int = 2, gen: torch.Generator | None = None,
                    device="cpu", channels: int | None = None) -> torch.Tensor:
    """On the CPU the draws come from ``gen``; on a GPU device (grids of up
    to 12 GB at 384^3 / SH 3) they come from a generator of that device
    seeded by one draw from ``gen``, and are made in place there."""
    C = channels_for_degree(sh_degree) if channels is None else channels
    V = (R + 1) ** 3
    ncoef_active = (sh_degree + 1) ** 2
    ncoef_slots = (C - 5) // 3
    assert ncoef_active <= ncoef_slots
    dev = torch.device(device)
    if dev.type != "cpu":
        seed = int(torch.randint(0, 2 ** 62, (1,), generator=gen, dtype=torch.int64))
        g = torch.Generator(device=dev).manual_seed(seed)
        out = torch.zeros((V, C), dtype=torch.float32, device=dev)
        out[:, :ncoef_active * 3].normal_(0.0, 0.1, generator=g)
        out[:, C - 5:].uniform_(0.0, 1.0, generator=g)
        return out
    out = torch.zeros((V, C), dtype=torch.float32)
    coef = torch.randn((V, ncoef_active * 3), generator=gen, dtype=torch.float32) * 0.1
    out[:, :ncoef_active * 3] = coef
    rest = torch.rand((V, 5), generator=gen, dtype=torch.float32)
    out[:, C - 5:] = rest
    return out


@dataclass
class Workload:
    name: str
    R: int
    sh_degree: int
    width: int
    height: int
    n_samples: int
    n_poses: int
    scene: str
    noise_world: float
    s: float = 5.0
    extra: dict = field(default_factory=dict)

    @property
    def interval(self) -> float:
        return interval_for(self.n_samples)


WORKLOADS = {
    # sphere r=1 + N(0,0.01), SH degree 1, 100x100, N=128, one pose
    "small": Workload("sphere64_100x100_n128", 64, 1, 100, 100, 128, 1, "sphere", 0.01),
    # torus R=1 r=0.35 + N(0,0.005), SH degree 2, 800x800, N=192, 100 poses
    "plain": Workload("torus128_800x800_n192", 128, 2, 800, 800, 192, 100, "torus", 0.005),
    # union of 8 random spheres, SH degree 2, 65,536 random rays/step, N=256
    "train": Workload("spheres256_65536rays_n256", 256, 2, 800, 800, 256, 100, "spheres", 0.0),
    # gyroid shell, SH degree 3, 1920x1080, N=384
    "large": Workload("gyroid384_1920x1080_n384", 384, 3, 1920, 1080, 384, 50, "gyroid", 0.005),
}


def scene_phi(name: str, seed: int = 0):
    if name == "spheres":
        return make_sphere_union(seed)
    return SCENES[name]


def build_grids(w: Workload, seed: int = 0, device="cpu"):
    """(sdf_grid [R,R,R,7], data_grid [(R+1)^3, C], generator) for a
    workload; the generator is returned so the loss target comes next in the
    fixed draw order."""
    gen = torch.Generator().manual_seed(seed)
    grid = quadric_grid(scene_phi(w.scene, seed), w.R, w.noise_world, gen, device=device)
    data = appearance_grid(w.R, w.sh_degree, gen, device=device)
    return grid, data, gen
