"""Host-side validation of the drop-in ``qsdvr`` module: bad tensors raise
ValueError before any GPU work (runs without a GPU)."""
import numpy as np
import pytest
import torch

from qsdvr_amd import qsdvr


def test_rejects_cpu_and_wrong_dtype():
    grid = torch.zeros((4, 4, 4, 7))
    pts = torch.zeros((3, 3))
    idx = torch.zeros(3, dtype=torch.int32)
    out = torch.zeros(3)
    with pytest.raises(ValueError, match="cuda"):
        qsdvr.SampleSDFForward(out, grid, pts, idx, 4)
    with pytest.raises(ValueError, match="dtype"):
        qsdvr.SampleSDFForward(out, grid.double(), pts, idx, 4)
    with pytest.raises(ValueError, match="dtype"):
        qsdvr.GridInterpolationForward(grid.double(), pts, idx, out, 5)


def test_rejects_shapes_and_layout():
    t = torch.zeros((6, 3)).t()
    with pytest.raises(ValueError, match="contiguous"):
        qsdvr.ShaderForward(torch.zeros((3, 3)), t, t, torch.zeros((3, 32)), torch.zeros((3, 3)))
    with pytest.raises(ValueError):
        qsdvr.RayAggregateForward(torch.zeros((2, 3)), torch.zeros((4, 3)), torch.zeros(5),
                                  torch.zeros((2, 3), dtype=torch.int32), torch.zeros(4),
                                  torch.zeros(4), 5.0)
    with pytest.raises(ValueError, match="7"):
        qsdvr.SampleNormalForward(torch.zeros((1, 3)), torch.zeros(13), torch.zeros((1, 3)),
                                  torch.zeros(1, dtype=torch.int32))


def test_camera_args_mirror():
    a = qsdvr.CameraArgs()
    a.width, a.height = 4, 3
    a.rotation = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    a.translation = [0.0, 0.0, -3.0]
    assert a.rotation.dtype == np.float32 and a.rotation.shape == (9,)
    s = a.to_struct()
    assert s.width == 4 and list(s.translation) == [0.0, 0.0, -3.0]
    a.pixelCount = 11
    with pytest.raises(ValueError, match="pixelCount"):
        qsdvr.RayCut(a, None, torch.zeros((12, 4), dtype=torch.int32), torch.zeros((12, 3)),
                     torch.zeros((12, 3)), 0.1)
    a.pixelCount = 12
    with pytest.raises(ValueError, match="cuda"):
        qsdvr.RayCut(a, None, torch.zeros((12, 4), dtype=torch.int32), torch.zeros((12, 3)),
                     torch.zeros((12, 3)), 0.1)


def test_interpolation_backward_path_choice(monkeypatch):
    """GridInterpolationBackward takes the sorted scatter wherever its keys
    (parity class and cell, up to R^3 < 2^31) fit at an even resolution, and
    the atomic kernel (no size limit) everywhere else -- never an error at
    sizes the reference runs."""
    n = 1000
    assert qsdvr.SCATTER == "sorted"
    assert qsdvr.interp_backward_path(n, 129, 32) == "sorted"
    assert qsdvr.interp_backward_path(n, 385, 53) == "sorted"
    assert qsdvr.interp_backward_path(n, 513, 32) == "sorted"    # 2^27 cells
    assert qsdvr.interp_backward_path(n, 1291, 32) == "sorted"   # 1290^3 < 2^31 - 1 cells
    assert qsdvr.interp_backward_path(n, 1293, 32) == "atomic"   # 1292^3 > 2^31 - 1 cells
    assert qsdvr.interp_backward_path(n, 128, 32) == "atomic"    # odd cell resolution
    assert qsdvr.interp_backward_path(n, 129, 40) == "atomic"    # channel count
    assert qsdvr.interp_backward_path(0, 129, 32) == "atomic"
    assert qsdvr.interp_backward_path(1 << 32, 129, 32) == "atomic"
    monkeypatch.setattr(qsdvr, "SCATTER", "atomic")
    assert qsdvr.interp_backward_path(n, 129, 32) == "atomic"


def test_quadric_backward_path_choice(monkeypatch):
    g = torch.zeros((16, 16, 16, 7))
    assert qsdvr.quadric_backward_path(100, g) == "sorted"
    assert qsdvr.quadric_backward_path(0, g) == "atomic"
    assert qsdvr.quadric_backward_path(1 << 32, g) == "atomic"
    monkeypatch.setattr(qsdvr, "SCATTER", "atomic")
    assert qsdvr.quadric_backward_path(100, g) == "atomic"
