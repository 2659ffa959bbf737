"""Ray generation on the MI355X vs the bit-exact float32 oracle
(kernel/camera.cu): RayInfo counts/offsets, totals, cell indices and
sample points must match bit for bit."""
import numpy as np
import pytest
import torch

from oracle import raygen
from qsdvr_amd import qsdvr, scenes
from qsdvr_amd.camera import Camera

pytestmark = pytest.mark.gpu


def _gpu_cut(cam, interval, mask=None):
    a = qsdvr.CameraArgs.from_camera(cam)
    P = a.pixelCount
    ray = torch.zeros((P, 4), dtype=torch.int32, device="cuda")
    ori = torch.zeros((P, 3), dtype=torch.float32, device="cuda")
    dirs = torch.zeros((P, 3), dtype=torch.float32, device="cuda")
    qsdvr.RayCut(a, mask, ray, ori, dirs, interval)
    torch.cuda.synchronize()
    return a, ray.cpu().numpy(), ori.cpu().numpy(), dirs.cpu().numpy()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("W,H,R,N,pose", [(100, 100, 64, 128, 0), (37, 23, 16, 40, 3),
                                          (64, 48, 32, 96, 7)])
def test_ray_cut_and_points_bit_exact(W, H, R, N, pose):
    cam = scenes.orbit_cameras(8, W, H)[pose]
    interval = np.float32(scenes.interval_for(N))
    a, ray, ori, dirs = _gpu_cut(cam, float(interval))
    ray_o, ori_o, dir_o, st, rt = raygen.ray_cut(cam.as_dict(), None, interval)
    assert np.array_equal(ray, ray_o)
    assert (a.sdfCount, a.renderCount) == (st, rt)
    assert np.array_equal(_bits(dirs), _bits(dir_o))
    assert np.array_equal(_bits(ori), _bits(ori_o))
    c = Camera(a, float(interval), R)
    c.GenerateRayPoints(None)
    d = c.GetRayPoints()
    torch.cuda.synchronize()
    sp, si, rp, ri, vd = raygen.generate_ray_points(ray_o, ori_o, dir_o, interval, R)
    assert np.array_equal(d["sdfIndexList"].cpu().numpy(), si)
    assert np.array_equal(d["renderIndexList"].cpu().numpy(), ri)
    assert np.array_equal(_bits(d["sdfPointList"].cpu().numpy()), _bits(sp))
    assert np.array_equal(_bits(d["renderPointList"].cpu().numpy()), _bits(rp))
    assert np.array_equal(_bits(d["viewDirList"].cpu().numpy()), _bits(vd))


def test_ray_cut_with_mask_bit_exact():
    cam = scenes.orbit_cameras(4, 60, 40)[1]
    interval = np.float32(scenes.interval_for(64))
    g = np.random.default_rng(3)
    mask = g.random(60 * 40) < 0.6
    a, ray, ori, dirs = _gpu_cut(cam, float(interval), torch.as_tensor(mask))
    ray_o, ori_o, dir_o, st, rt = raygen.ray_cut(cam.as_dict(), mask, interval)
    # masked rows are fully defined; unmasked rows carry the raw count only
    assert np.array_equal(ray[mask], ray_o[mask])
    assert np.array_equal(ray[~mask][:, 0], ray_o[~mask][:, 0])
    assert (a.sdfCount, a.renderCount) == (st, rt)


def test_full_frame_pose_bit_exact():
    """One 800x800 pose at R=128, N=192: every RayInfo row and total; sample
    lists compared on a strided subset of rays."""
    cam = scenes.orbit_cameras(100, 800, 800)[17]
    interval = np.float32(scenes.interval_for(192))
    a, ray, ori, dirs = _gpu_cut(cam, float(interval))
    ray_o, ori_o, dir_o, st, rt = raygen.ray_cut(cam.as_dict(), None, interval)
    assert np.array_equal(ray, ray_o)
    assert (a.sdfCount, a.renderCount) == (st, rt)
    assert np.array_equal(_bits(ori), _bits(ori_o))
    c = Camera(a, float(interval), 128)
    c.GenerateRayPoints(None)
    d = c.GetRayPoints()
    sub = np.arange(0, ray_o.shape[0], 997)
    sp, si, rp, ri, vd = raygen.generate_ray_points(ray_o[sub], ori_o[sub], dir_o[sub], interval,
                                                    128)
    rows = ray_o[sub].view(np.uint32).astype(np.int64)
    sidx = np.concatenate([np.arange(o, o + n) for n, o in rows[:, :2]])
    ridx = np.concatenate([np.arange(o, o + n) for n, o in rows[:, 2:]])
    assert np.array_equal(d["sdfIndexList"].cpu().numpy()[sidx], si[sidx])
    assert np.array_equal(_bits(d["sdfPointList"].cpu().numpy()[sidx]), _bits(sp[sidx]))
    assert np.array_equal(d["renderIndexList"].cpu().numpy()[ridx], ri[ridx])
    assert np.array_equal(_bits(d["renderPointList"].cpu().numpy()[ridx]), _bits(rp[ridx]))
