"""Drop-in replacement of the reference extension module ``qsdvr``.

Same 14 functions, argument order, in-place output conventions and
``CameraArgs`` attributes as kernel/qsdvr.cpp:5-50, so the reference's
callers (models/render.py, datasets/utils.py) run unchanged with
``import qsdvr_amd.qsdvr as qsdvr``.  Each function validates its tensors
(dtype, device, contiguity, shapes) and raises ``ValueError`` on bad input,
then calls the HIP C-ABI (include/qsdvr_hip.h) on torch's current stream;
a nonzero status raises ``qsdvr_amd._lib.QsdvrError``.

Output conventions (kernel/qsdvr.cpp, models/render.py):
  * forward ops overwrite their outputs;
  * GridInterpolationBackward, SampleSDFBackward and SampleNormalBackward
    accumulate into a caller-zeroed grid;
  * ShaderBackward, RayAggregateBackward and LayerToGridBackward overwrite.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib

__all__ = ["CameraArgs", "GridInterpolationForward", "GridInterpolationBackward",
           "LayerToGridForward", "LayerToGridBackward", "SampleSDFForward",
           "SampleSDFBackward", "SampleNormalForward", "SampleNormalBackward",
           "ShaderForward", "ShaderBackward", "RayAggregateForward", "RayAggregateBackward",
           "RayCut", "GenerateRayPoints"]


class CameraArgs:
    """Mirror of the bound ``CameraArgs`` class (kernel/qsdvr.cpp:36-50,
    include/camera.h:19-34).  ``rotation`` (9) and ``translation`` (3) are
    float32 arrays assignable from lists; ``sdfCount`` / ``renderCount`` are
    written by ``RayCut``."""

    def __init__(self):
        self.width = 0
        self.height = 0
        self._rotation = np.zeros(9, dtype=np.float32)
        self._translation = np.zeros(3, dtype=np.float32)
        self.fx = 0.0
        self.fy = 0.0
        self.cx = 0.0
        self.cy = 0.0
        self.near = 0.0
        self.far = 0.0
        self.pixelCount = 0
        self.sdfCount = 0
        self.renderCount = 0

    @property
    def rotation(self):
        return self._rotation

    @rotation.setter
    def rotation(self, v):
        self._rotation = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))

    @property
    def translation(self):
        return self._translation

    @translation.setter
    def translation(self, v):
        self._translation = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))

    def to_struct(self) -> _lib.CameraStruct:
        if self._rotation.size != 9 or self._translation.size != 3:
            raise ValueError("CameraArgs: rotation needs 9 floats and translation 3")
        s = _lib.CameraStruct()
        s.width = int(self.width)
        s.height = int(self.height)
        for i in range(9):
            s.rotation[i] = float(self._rotation[i])
        for i in range(3):
            s.translation[i] = float(self._translation[i])
        s.fx, s.fy, s.cx, s.cy = float(self.fx), float(self.fy), float(self.cx), float(self.cy)
        s.near_depth, s.far_depth = float(self.near), float(self.far)
        return s

    @classmethod
    def from_camera(cls, cam) -> "CameraArgs":
        """From a ``qsdvr_amd.scenes.Camera`` (or any object with the same
        fields)."""
        a = cls()
        a.width, a.height = int(cam.width), int(cam.height)
        a.rotation = cam.rotation
        a.translation = cam.translation
        a.fx, a.fy, a.cx, a.cy = cam.fx, cam.fy, cam.cx, cam.cy
        a.near, a.far = cam.near, cam.far
        a.pixelCount = a.width * a.height
        return a


# ------------------------------------------------------------ checking ----
def _t(name, t, dtype, device="cuda", ndim=None, shape=None):
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch.Tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: expected dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    if ndim is not None and t.dim() != ndim:
        raise ValueError(f"{name}: expected {ndim} dims, got shape {tuple(t.shape)}")
    if shape is not None:
        if t.dim() < len(shape):
            raise ValueError(f"{name}: expected shape {shape}, got {tuple(t.shape)}")
        for i, s in enumerate(shape):
            if s is not None and t.shape[i] != s:
                raise ValueError(f"{name}: expected shape {shape}, got {tuple(t.shape)}")
    if device is not None and t.device.type != device:
        raise ValueError(f"{name}: expected a {device} tensor, got {t.device}")
    return t


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else C.c_void_p(0)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _points(name, t, n=None):
    return _t(name, t, torch.float32, ndim=2, shape=(n, 3))


def _channels(grid):
    return int(grid.shape[-1])


# Every drop-in op counts its calls, so that GridInterpolationBackward's
# sort is reused only by a SampleNormalBackward that immediately follows it
# (see _remember_render_sort).
_calls = 0


def _op(fn):
    import functools

    @functools.wraps(fn)
    def wrapped(*args, **kwargs):
        global _calls, _render_sort, _forward_sort
        _calls += 1
        result = fn(*args, **kwargs)
        if _render_sort is not None and _render_sort[-1] != _calls:
            _render_sort = None  # an earlier call's sort: no later call may use it
        if _forward_sort is not None and _forward_sort[0]() is None:
            _forward_sort = None  # its samples are gone: free the workspace
        return result
    return wrapped


# ------------------------------------------------------ render ops ----
@_op
def GridInterpolationForward(dataGrid, pointList, indexList, out, reso):
    """kernel/render.cu:317-323; oracle semantics of render_torch.py:128-143.
    ``reso`` is dataGrid.size(0) = R+1, as models/render.py:15 passes it."""
    _t("dataGrid", dataGrid, torch.float32)
    n = _points("pointList", pointList).shape[0]
    _t("indexList", indexList, torch.int32, shape=(n,))
    Cc = _channels(dataGrid)
    _t("out", out, torch.float32, shape=(n, Cc))
    R1 = int(reso)
    if dataGrid.numel() != R1 ** 3 * Cc:
        raise ValueError(f"dataGrid has {dataGrid.numel()} elements, expected ({R1})^3 x {Cc}")
    if interp_forward_path(n, R1, Cc, dataGrid, out) == "sorted":
        ws_n = _sorted_workspace(n, (R1 - 1) ** 3)
        ws = _workspace(ws_n)
        _lib.call("qsdvr_grid_interpolation_forward_sorted", _ptr(dataGrid), _ptr(pointList),
                  _ptr(indexList), _ptr(out), n, R1, Cc, _ptr(ws), ws_n, _stream())
        _remember_forward_sort(pointList, indexList, n, R1 - 1, ws, ws_n)
        return
    _lib.call("qsdvr_grid_interpolation_forward", _ptr(dataGrid), _ptr(pointList),
              _ptr(indexList), _ptr(out), n, R1, Cc, _stream())


@_op
def GridInterpolationBackward(dataGrid, pointList, indexList, out, reso):
    """kernel/render.cu:325-331.  Roles as in the reference: ``dataGrid`` is
    the zeroed gradient accumulator, ``out`` the upstream gradient."""
    _t("dataGrid", dataGrid, torch.float32)
    n = _points("pointList", pointList).shape[0]
    _t("indexList", indexList, torch.int32, shape=(n,))
    Cc = _channels(dataGrid)
    _t("out", out, torch.float32, shape=(n, Cc))
    R1 = int(reso)
    if dataGrid.numel() != R1 ** 3 * Cc:
        raise ValueError(f"dataGrid has {dataGrid.numel()} elements, expected ({R1})^3 x {Cc}")
    if interp_backward_path(n, R1, Cc) == "sorted":
        fwd = _take_forward_sort(pointList, indexList, n, R1 - 1)
        if fwd is not None:  # GridInterpolationForward sorted these samples
            ws, ws_n = fwd
            _lib.call("qsdvr_grid_interpolation_backward_presorted", _ptr(dataGrid), _ptr(out),
                      n, R1, Cc, _ptr(ws), ws_n, _stream())
            _remember_render_sort(pointList, indexList, n, R1 - 1, ws, ws_n)
            return
        ws_n = _sorted_workspace(n, (R1 - 1) ** 3)
        ws = _workspace(ws_n)
        _lib.call("qsdvr_grid_interpolation_backward_sorted", _ptr(dataGrid), _ptr(pointList),
                  _ptr(indexList), _ptr(out), n, R1, Cc, _ptr(ws), ws_n, _stream())
        _remember_render_sort(pointList, indexList, n, R1 - 1, ws, ws_n)
        return
    _lib.call("qsdvr_grid_interpolation_backward", _ptr(dataGrid), _ptr(pointList),
              _ptr(indexList), _ptr(out), n, R1, Cc, _stream())


@_op
def ShaderForward(out, normalList, viewDirList, dataList, specularList):
    """kernel/render.cu:333-339; render_torch.py:144-152."""
    n = _points("normalList", normalList).shape[0]
    _points("viewDirList", viewDirList, n)
    _t("dataList", dataList, torch.float32, ndim=2, shape=(n, None))
    _points("out", out, n)
    _points("specularList", specularList, n)
    _lib.call("qsdvr_shader_forward", _ptr(out), _ptr(normalList), _ptr(viewDirList),
              _ptr(dataList), _ptr(specularList), n, int(dataList.shape[1]), _stream())


@_op
def ShaderBackward(out, normalList, viewDirList, dataList, specularList, normalGradList,
                   dataGradList):
    """kernel/render.cu:341-347.  ``out`` is the upstream gradient."""
    n = _points("normalList", normalList).shape[0]
    _points("out", out, n)
    _points("viewDirList", viewDirList, n)
    _t("dataList", dataList, torch.float32, ndim=2, shape=(n, None))
    _points("specularList", specularList, n)
    _points("normalGradList", normalGradList, n)
    _t("dataGradList", dataGradList, torch.float32, shape=tuple(dataList.shape))
    _lib.call("qsdvr_shader_backward", _ptr(out), _ptr(normalList), _ptr(viewDirList),
              _ptr(dataList), _ptr(specularList), _ptr(normalGradList), _ptr(dataGradList), n,
              int(dataList.shape[1]), _stream())


def _raylist(rayList):
    return _t("rayList", rayList, torch.int32, ndim=2, shape=(None, 4)).shape[0]


@_op
def RayAggregateForward(out, rgbList, sdfList, rayList, alphaList, TList, logisticCoef):
    """kernel/render.cu:349-355; render_torch.py:155-167."""
    rays = _raylist(rayList)
    nr = _points("rgbList", rgbList).shape[0]
    ns = _t("sdfList", sdfList, torch.float32, ndim=1).shape[0]
    _t("out", out, torch.float32, shape=(rays, 3))
    _t("alphaList", alphaList, torch.float32, shape=(nr,))
    _t("TList", TList, torch.float32, shape=(nr,))
    _lib.call("qsdvr_ray_aggregate_forward", _ptr(out), _ptr(rgbList), _ptr(sdfList),
              _ptr(rayList), _ptr(alphaList), _ptr(TList), rays, ns, nr, float(logisticCoef),
              _stream())


@_op
def RayAggregateBackward(outgrad, rgbList, sdfList, rayList, alphaList, TList, rgbGradList,
                         sdfGradList, logisticCoef):
    """kernel/render.cu:357-363."""
    rays = _raylist(rayList)
    nr = _points("rgbList", rgbList).shape[0]
    ns = _t("sdfList", sdfList, torch.float32, ndim=1).shape[0]
    _t("outgrad", outgrad, torch.float32, shape=(rays, 3))
    _t("alphaList", alphaList, torch.float32, shape=(nr,))
    _t("TList", TList, torch.float32, shape=(nr,))
    _points("rgbGradList", rgbGradList, nr)
    _t("sdfGradList", sdfGradList, torch.float32, shape=(ns,))
    _lib.call("qsdvr_ray_aggregate_backward", _ptr(outgrad), _ptr(rgbList), _ptr(sdfList),
              _ptr(rayList), _ptr(alphaList), _ptr(TList), _ptr(rgbGradList),
              _ptr(sdfGradList), rays, ns, nr, float(logisticCoef), _stream())


# ------------------------------------------------------ surface ops ----
def _grid7(name, g):
    _t(name, g, torch.float32, device=None)
    if g.numel() % 7 != 0 or g.numel() == 0:
        raise ValueError(f"{name}: numel {g.numel()} is not a positive multiple of 7")
    _t(name, g, torch.float32)
    return g.numel() // 7


@_op
def SampleSDFForward(out, sdfGrid, pointList, indexList, reso):
    """kernel/surface.cu:343-350; render_torch.py:77-90."""
    cells = _grid7("sdfGrid", sdfGrid)
    n = _points("pointList", pointList).shape[0]
    _t("indexList", indexList, torch.int32, shape=(n,))
    _t("out", out, torch.float32, shape=(n,))
    if float(reso) == 0.0:
        raise ValueError("reso must be nonzero")
    _lib.call("qsdvr_sample_sdf_forward", _ptr(out), _ptr(sdfGrid), _ptr(pointList),
              _ptr(indexList), n, cells, float(reso), _stream())


@_op
def SampleSDFBackward(out, sdfGrid, pointList, indexList, sdfGradGrid, reso):
    """kernel/surface.cu:352-359.  ``out`` is the upstream gradient."""
    cells = _grid7("sdfGrid", sdfGrid)
    n = _points("pointList", pointList).shape[0]
    _t("indexList", indexList, torch.int32, shape=(n,))
    _t("out", out, torch.float32, shape=(n,))
    _t("sdfGradGrid", sdfGradGrid, torch.float32)
    if sdfGradGrid.numel() != sdfGrid.numel():
        raise ValueError("sdfGradGrid must match sdfGrid")
    if float(reso) == 0.0:
        raise ValueError("reso must be nonzero")
    if quadric_backward_path(n, sdfGrid) == "sorted":
        ws_n = _sorted_workspace(n, cells)
        ws = _workspace(ws_n)
        _lib.call("qsdvr_sample_sdf_backward_sorted", _ptr(out), _ptr(sdfGrid), _ptr(pointList),
                  _ptr(indexList), _ptr(sdfGradGrid), n, cells, float(reso), _ptr(ws), ws_n,
                  _stream())
        return
    _lib.call("qsdvr_sample_sdf_backward", _ptr(out), _ptr(sdfGrid), _ptr(pointList),
              _ptr(indexList), _ptr(sdfGradGrid), n, cells, float(reso), _stream())


@_op
def SampleNormalForward(out, sdfGrid, pointList, indexList):
    """kernel/surface.cu:361-367; render_torch.py:91-100."""
    cells = _grid7("sdfGrid", sdfGrid)
    n = _points("pointList", pointList).shape[0]
    _t("indexList", indexList, torch.int32, shape=(n,))
    _points("out", out, n)
    _lib.call("qsdvr_sample_normal_forward", _ptr(out), _ptr(sdfGrid), _ptr(pointList),
              _ptr(indexList), n, cells, _stream())


@_op
def SampleNormalBackward(out, sdfGrid, pointList, indexList, sdfGradGrid):
    """kernel/surface.cu:369-375.  ``out`` is the upstream gradient."""
    cells = _grid7("sdfGrid", sdfGrid)
    n = _points("pointList", pointList).shape[0]
    _t("indexList", indexList, torch.int32, shape=(n,))
    _points("out", out, n)
    _t("sdfGradGrid", sdfGradGrid, torch.float32)
    if sdfGradGrid.numel() != sdfGrid.numel():
        raise ValueError("sdfGradGrid must match sdfGrid")
    if quadric_backward_path(n, sdfGrid) == "sorted":
        reuse = _take_render_sort(pointList, indexList, n, cells)
        if reuse is not None:  # GridInterpolationBackward already sorted these samples
            R, ws, ws_n = reuse
            _lib.call("qsdvr_sample_normal_backward_presorted", _ptr(out), _ptr(sdfGrid),
                      _ptr(sdfGradGrid), n, R, _ptr(ws), ws_n, _stream())
            return
        ws_n = _sorted_workspace(n, cells)
        ws = _workspace(ws_n)
        _lib.call("qsdvr_sample_normal_backward_sorted", _ptr(out), _ptr(sdfGrid),
                  _ptr(pointList), _ptr(indexList), _ptr(sdfGradGrid), n, cells, _ptr(ws), ws_n,
                  _stream())
        return
    _lib.call("qsdvr_sample_normal_backward", _ptr(out), _ptr(sdfGrid), _ptr(pointList),
              _ptr(indexList), _ptr(sdfGradGrid), n, cells, _stream())


def _layers(x, y, z, off):
    dev = x.device.type
    for nm, t in (("xLayer", x), ("yLayer", y), ("zLayer", z), ("offset", off)):
        _t(nm, t, torch.float32, device=None, ndim=1)
        if t.device.type != dev:
            raise ValueError("layers and offset must share one device")
    R = x.shape[0]
    if y.shape[0] != R or z.shape[0] != R or off.shape[0] != 4:
        raise ValueError("layers must have equal length R and offset 4 entries")
    return R, dev


def _workspace(nbytes):
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device="cuda")


# How the three accumulating backward ops (GridInterpolationBackward,
# SampleSDFBackward, SampleNormalBackward) add into the gradient grid; both
# give the sums of the reference's one float atomic per sample, coefficient
# and channel (kernel/render.cu:121-152, kernel/surface.cu:74-146):
#   "sorted": sorted-segment sums in a fixed order -- bitwise reproducible
#             and several times faster on per-pose sample lists, where the
#             samples of many rays pile onto the same cells;
#   "atomic": one atomic per sample, coefficient and channel.
# Sizes the sorted form does not support fall back to "atomic".
SCATTER = "sorted"


def _sorted_fits(n, cells):
    """The sorted-segment kernels key samples by cell in 32 bits and count
    them in 32 bits."""
    return 0 < n < 0xFFFFFFFF and 0 < cells < 0x7FFFFFFF


def interp_backward_path(n, vertex_reso, channels):
    """"sorted" or "atomic": which GridInterpolationBackward kernel runs for
    ``SCATTER``.  The sorted form orders the samples by cell parity class
    and cell (keys up to R^3) and needs an even cell resolution (cells of
    one parity class then share no vertex row) and 32 or 53 channels."""
    R = int(vertex_reso) - 1
    ok = (SCATTER == "sorted" and _sorted_fits(n, R ** 3) and channels in (32, 53) and R >= 2
          and R % 2 == 0)
    return "sorted" if ok else "atomic"


def interp_forward_path(n, vertex_reso, channels, dataGrid=None, out=None):
    """"sorted" or "gather": which GridInterpolationForward kernel runs.  The
    sorted form makes GridInterpolationBackward's sort of the samples (so the
    backward on the same lists skips it) and interpolates one cell's run of
    samples at a time: the sorted backward's sizes, 32 channels, and 16-B
    aligned grid and output rows."""
    aligned = all(t is None or t.data_ptr() % 16 == 0 for t in (dataGrid, out))
    ok = interp_backward_path(n, vertex_reso, channels) == "sorted" and channels == 32 and aligned
    return "sorted" if ok else "gather"


def quadric_backward_path(n, sdfGrid):
    """"sorted" or "atomic": which SampleSDF / SampleNormal backward kernel
    runs for ``SCATTER``."""
    ok = SCATTER == "sorted" and _sorted_fits(n, sdfGrid.numel() // 7)
    return "sorted" if ok else "atomic"


# The reference's backward (models/render.py) runs GridInterpolationBackward
# and then SampleNormalBackward on the same render point and index lists.
# The sorted GridInterpolationBackward orders those samples (point and index)
# by (parity class, cell), which also groups equal cells, so
# SampleNormalBackward can sum its runs on that order instead of sorting
# again.  The workspace holding it is kept here (one entry) with weak
# references to the point and index tensors, their version counters (an
# in-place change invalidates it), the size and resolution, the stream and
# the op-call count: only the very next op call, if it is
# SampleNormalBackward on those tensors, takes it, and any other next call
# drops it (freeing the workspace).  (The reused order sums a
# cell's samples in a different tree than SampleNormalBackward's own sort,
# so the bits then depend on the preceding call -- the same sequence of
# calls still gives the same bits.)
_render_sort = None


def _remember_render_sort(points, index, n, R, ws, ws_n):
    global _render_sort
    import weakref
    _render_sort = (weakref.ref(points), points._version, weakref.ref(index), index._version,
                    int(n), int(R), ws, int(ws_n), torch.cuda.current_stream().cuda_stream,
                    _calls)


def _take_render_sort(points, index, n, cells):
    """(R, workspace, bytes) of a matching GridInterpolationBackward sort, or
    None; the entry is consumed."""
    global _render_sort
    ent, _render_sort = _render_sort, None
    if ent is None:
        return None
    pref, pversion, iref, iversion, n0, R, ws, ws_n, stream, call = ent
    if (call + 1 != _calls or pref() is not points or points._version != pversion
            or iref() is not index or index._version != iversion or n0 != int(n)
            or R ** 3 != int(cells) or stream != torch.cuda.current_stream().cuda_stream):
        return None
    return R, ws, ws_n


# GridInterpolationForward's sort, for the GridInterpolationBackward of the
# same samples (models/render.py saves the point and index lists for it):
# one entry, with weak references to the point and index tensors and their
# version counters (an in-place change invalidates it), the size and
# resolution and the stream.  The matching backward takes it; the next
# sorted forward replaces it; it is dropped when its point list is freed.
# Between forward and backward it holds the workspace (about 41 B per sample
# plus the sort's temporary storage).
_forward_sort = None


def _remember_forward_sort(points, index, n, R, ws, ws_n):
    global _forward_sort
    import weakref
    _forward_sort = (weakref.ref(points), points._version, weakref.ref(index), index._version,
                     int(n), int(R), ws, int(ws_n), torch.cuda.current_stream().cuda_stream)


def _take_forward_sort(points, index, n, R):
    """(workspace, bytes) of a matching GridInterpolationForward sort, or None;
    the entry is consumed either way."""
    global _forward_sort
    ent, _forward_sort = _forward_sort, None
    if ent is None:
        return None
    pref, pversion, iref, iversion, n0, R0, ws, ws_n, stream = ent
    if (pref() is not points or points._version != pversion or iref() is not index
            or index._version != iversion or n0 != int(n) or R0 != int(R)
            or stream != torch.cuda.current_stream().cuda_stream):
        return None
    return ws, ws_n


def _sorted_workspace(n, cells):
    nb = int(_lib.fn("qsdvr_sorted_scatter_workspace_size")(int(n), int(cells)))
    if nb == 0:
        raise ValueError(f"{n} samples / {cells} cells exceed the sorted scatter's 32-bit keys")
    return nb


@_op
def LayerToGridForward(out, xLayer, yLayer, zLayer, offset):
    """kernel/surface.cu:148-229; render_torch.py:30-75.  Layers may be CPU
    tensors (as in models/render.py:177-183) or CUDA tensors."""
    R, dev = _layers(xLayer, yLayer, zLayer, offset)
    _t("out", out, torch.float32)
    if out.numel() != R ** 3 * 7:
        raise ValueError(f"out must hold {R}^3 x 7 floats")
    ws_n = _lib.fn("qsdvr_layer_to_grid_workspace_size")(R)
    ws = _workspace(ws_n)
    _lib.call("qsdvr_layer_to_grid_forward", _ptr(out), _ptr(xLayer), _ptr(yLayer),
              _ptr(zLayer), _ptr(offset), R, 1 if dev == "cpu" else 0, _ptr(ws), ws_n, _stream())


@_op
def LayerToGridBackward(out, xLayer, yLayer, zLayer, offset):
    """kernel/surface.cu:231-341.  ``out`` is the upstream gradient grid; the
    layer tensors and offset are OUTPUTS (CPU as in models/render.py:117-120,
    or CUDA), overwritten.  offset[3] gets the sum of all G gradients."""
    R, dev = _layers(xLayer, yLayer, zLayer, offset)
    _t("out", out, torch.float32)
    if out.numel() != R ** 3 * 7:
        raise ValueError(f"out must hold {R}^3 x 7 floats")
    ws_n = _lib.fn("qsdvr_layer_to_grid_workspace_size")(R)
    ws = _workspace(ws_n)
    _lib.call("qsdvr_layer_to_grid_backward", _ptr(out), _ptr(xLayer), _ptr(yLayer),
              _ptr(zLayer), _ptr(offset), R, 1 if dev == "cpu" else 0, _ptr(ws), ws_n,
              _stream())


# ------------------------------------------------------- camera ops ----
@_op
def RayCut(args: CameraArgs, mask, rayList, originList, dirList, interval):
    """kernel/camera.cu:98-130.  ``mask``: bool [pixelCount] (CPU as in the
    reference; a CUDA mask is copied to the host).  Sets ``args.sdfCount``
    and ``args.renderCount``."""
    P = int(args.pixelCount)
    if P != int(args.width) * int(args.height):
        raise ValueError(f"pixelCount {P} != width*height")
    _t("rayList", rayList, torch.int32, shape=(P, 4))
    _t("originList", originList, torch.float32, shape=(P, 3))
    _t("dirList", dirList, torch.float32, shape=(P, 3))
    if float(interval) <= 0.0:
        raise ValueError("interval must be > 0")
    mask_np = None
    if mask is not None:
        if not isinstance(mask, torch.Tensor) or mask.dtype != torch.bool:
            raise ValueError("mask must be a bool tensor")
        if mask.numel() != P:
            raise ValueError(f"mask has {mask.numel()} entries, expected {P}")
        mask_np = np.ascontiguousarray(mask.detach().cpu().numpy().reshape(-1).astype(np.uint8))
    ws_n = _lib.fn("qsdvr_ray_cut_workspace_size")(P)
    ws = _workspace(ws_n)
    st, rt = C.c_int64(0), C.c_int64(0)
    cam = args.to_struct()
    mptr = C.c_void_p(mask_np.ctypes.data) if mask_np is not None else C.c_void_p(0)
    _lib.call("qsdvr_ray_cut", C.byref(cam), mptr, _ptr(rayList), _ptr(originList),
              _ptr(dirList), float(interval), P, _ptr(ws), ws_n, C.byref(st), C.byref(rt),
              _stream())
    args.sdfCount = int(st.value)
    args.renderCount = int(rt.value)


@_op
def GenerateRayPoints(args: CameraArgs, sdfPointList, sdfIndexList, renderPointList,
                      renderIndexList, viewDirList, rayList, originList, dirList, interval,
                      baseVolumeReso):
    """kernel/camera.cu:132-146.  One row of rayList per ray (masked rays)."""
    rays = _raylist(rayList)
    ns = _points("sdfPointList", sdfPointList).shape[0]
    _t("sdfIndexList", sdfIndexList, torch.int32, shape=(ns,))
    nr = _points("renderPointList", renderPointList).shape[0]
    _t("renderIndexList", renderIndexList, torch.int32, shape=(nr,))
    _points("viewDirList", viewDirList, nr)
    _points("originList", originList, rays)
    _points("dirList", dirList, rays)
    _lib.call("qsdvr_generate_ray_points", _ptr(sdfPointList), _ptr(sdfIndexList),
              _ptr(renderPointList), _ptr(renderIndexList), _ptr(viewDirList), _ptr(rayList),
              _ptr(originList), _ptr(dirList), rays, ns, nr, float(interval),
              int(baseVolumeReso), _stream())
