"""ctypes binding of ``libqsdvr_hip.so`` (declared in include/qsdvr_hip.h).

The library is the only compute path: if it is missing or fails to load,
every op raises -- there is no CPU or PyTorch fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libqsdvr_hip.so"

QSDVR_OK = 0
STATUS_NAMES = {-1: "EINVAL", -2: "EHIP", -3: "ERANGE", -4: "EWORKSPACE"}


class QsdvrError(RuntimeError):
    """A C-ABI call returned a nonzero status."""

    def __init__(self, name, status, message):
        super().__init__(f"{name} failed with {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class CameraStruct(C.Structure):
    """``qsdvr_camera`` (include/qsdvr_hip.h)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("rotation", C.c_float * 9), ("translation", C.c_float * 3),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("near_depth", C.c_float), ("far_depth", C.c_float)]


P = C.c_void_p
I64 = C.c_int64
I32 = C.c_int32
F = C.c_float
SZ = C.c_size_t

# name -> (argtypes, restype)
SIGNATURES = {
    "qsdvr_abi_version": ([], C.c_int),
    "qsdvr_last_error": ([], C.c_char_p),
    "qsdvr_ray_cut_workspace_size": ([I64], SZ),
    "qsdvr_ray_cut": ([C.POINTER(CameraStruct), P, P, P, P, F, I64, P, SZ,
                       C.POINTER(I64), C.POINTER(I64), P], C.c_int),
    "qsdvr_ray_cut_noalloc_max_pixels": ([], I64),
    "qsdvr_ray_cut_noalloc": ([C.POINTER(CameraStruct), P, P, P, P, F, I64,
                               C.POINTER(I64), C.POINTER(I64), P], C.c_int),
    "qsdvr_generate_ray_points": ([P, P, P, P, P, P, P, P, I64, I64, I64, F, I32, P], C.c_int),
    "qsdvr_layer_to_grid_workspace_size": ([I32], SZ),
    "qsdvr_layer_to_grid_forward": ([P, P, P, P, P, I32, C.c_int, P, SZ, P], C.c_int),
    "qsdvr_layer_to_grid_backward": ([P, P, P, P, P, I32, C.c_int, P, SZ, P], C.c_int),
    "qsdvr_sample_sdf_forward": ([P, P, P, P, I64, I64, F, P], C.c_int),
    "qsdvr_sample_sdf_backward": ([P, P, P, P, P, I64, I64, F, P], C.c_int),
    "qsdvr_sample_normal_forward": ([P, P, P, P, I64, I64, P], C.c_int),
    "qsdvr_sample_normal_backward": ([P, P, P, P, P, I64, I64, P], C.c_int),
    "qsdvr_grid_interpolation_forward": ([P, P, P, P, I64, I32, I32, P], C.c_int),
    "qsdvr_grid_interpolation_backward": ([P, P, P, P, I64, I32, I32, P], C.c_int),
    "qsdvr_shader_forward": ([P, P, P, P, P, I64, I32, P], C.c_int),
    "qsdvr_shader_backward": ([P, P, P, P, P, P, P, I64, I32, P], C.c_int),
    "qsdvr_ray_aggregate_forward": ([P, P, P, P, P, P, I64, I64, I64, F, P], C.c_int),
    "qsdvr_ray_aggregate_backward": ([P, P, P, P, P, P, P, P, I64, I64, I64, F, P], C.c_int),
    "qsdvr_sorted_scatter_workspace_size": ([I64, I64], SZ),
    "qsdvr_grid_interpolation_backward_sorted": ([P, P, P, P, I64, I32, I32, P, SZ, P], C.c_int),
    "qsdvr_sample_sdf_backward_sorted": ([P, P, P, P, P, I64, I64, F, P, SZ, P], C.c_int),
    "qsdvr_sample_normal_backward_sorted": ([P, P, P, P, P, I64, I64, P, SZ, P], C.c_int),
    "qsdvr_sample_normal_backward_presorted": ([P, P, P, I64, I32, P, SZ, P], C.c_int),
    "qsdvr_grid_interpolation_forward_sorted": ([P, P, P, P, I64, I32, I32, P, SZ, P], C.c_int),
    "qsdvr_grid_interpolation_backward_presorted": ([P, P, I64, I32, I32, P, SZ, P], C.c_int),
    "qsdvr_fused_forward": ([P, I32, I32, I32, P, I64, P, P, I32, I32, F, F, P, P, P],
                            C.c_int),
    "qsdvr_fused_backward": ([P, I32, I32, I32, P, I64, P, P, I32, I32, F, F, P, P, P, P, P,
                              P, I32, P], C.c_int),
    "qsdvr_fused_forward_bins": ([P, I32, I32, I32, P, I64, P, P, I32, I32, F, F, P, P, P, P],
                                 C.c_int),
    "qsdvr_ray_state": ([P, I32, I32, I32, P, I64, F, P, P], C.c_int),
    "qsdvr_plain_bins": ([P, I64, P, I32, F, F, P, P], C.c_int),
    "qsdvr_fused_order_workspace_size": ([I32, I32, I32], SZ),
    "qsdvr_fused_forward_costed": ([P, I32, I32, I32, P, P, I32, P, P, I32, I32, F, F, P, P, P, P],
                                   C.c_int),
    "qsdvr_fused_backward_ordered": ([P, I32, I32, I32, P, P, I32, P, P, I32, I32, F, F, P, P, P, P,
                                      P, P, P, P, SZ, P], C.c_int),
    "qsdvr_fused_forward_listed": ([P, I32, I32, I32, P, I64, P, P, I32, P, P, I32, I32, F, F,
                                    P, P, P], C.c_int),
    "qsdvr_fused_backward_listed": ([P, I32, I32, I32, P, I64, P, P, I32, P, P, I32, I32, F, F,
                                     P, P, P, P, P, P, I32, P], C.c_int),
    "qsdvr_fused_det_workspace_size": ([I64, I64, I32, I32], SZ),
    "qsdvr_fused_backward_det": ([P, I32, I32, I32, P, I64, P, P, I32, P, P, I32, I32, F, F,
                                  P, P, P, P, P, P, I64, P, SZ, P], C.c_int),
    "qsdvr_proposal_param_count": ([], C.c_int32),
    "qsdvr_proposal_packed_bytes": ([], SZ),
    "qsdvr_proposal_pack": ([P, P, P], C.c_int),
    "qsdvr_proposal_forward": ([P, I64, F, P, P, P, P, P, P, P], C.c_int),
    "qsdvr_proposal_workspace_size": ([I64], SZ),
    "qsdvr_proposal_backward": ([P, P, I64, P, P, P, P, P, P, SZ, P, P, P], C.c_int),
    "qsdvr_adam_step": ([P, P, P, P, I64, F, F, F, F, I32, P], C.c_int),
    "qsdvr_proposal_sample": ([P, P, I64, F, I32, F, F, P, P], C.c_int),
}

_lib = None
_path = None


def load(path: str | os.PathLike | None = None):
    """Load (once) and return the ctypes library; raises if it is missing.
    ``path`` loads another build of the same ABI instead of the in-tree
    library (``loaded_path()`` reports which one is in use)."""
    global _lib, _path
    if _lib is not None and path is None:
        return _lib
    # PyTorch ships its own HIP runtime under the same soname as ROCm's: it is
    # loaded first, so that the library's kernels register with the runtime
    # torch uses (a library loaded before torch pulls in ROCm's copy, and the
    # process then holds two runtimes, one of which finds no device).
    import torch  # noqa: F401
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"{p} is not built: run `python -m qsdvr_amd.build` (the HIP library is the only "
            "compute path; there is no fallback)")
    lib = C.CDLL(str(p))
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    _path = str(p.resolve())
    return lib


def loaded_path() -> str:
    """Absolute path of the library the ops call (loads it if needed)."""
    load()
    return _path


def call(name: str, *args) -> None:
    lib = load()
    fn = getattr(lib, name)
    st = fn(*args)
    if st != QSDVR_OK:
        msg = lib.qsdvr_last_error().decode(errors="replace")
        raise QsdvrError(name, st, msg)


def fn(name: str):
    return getattr(load(), name)


# The library is loaded when this module is first imported (by the op, fused,
# proposal and training modules), ahead of the caller's first GPU
# allocations: in the first process on a GPU, loading it only after the
# grids were allocated measured the fused adjoint 15-35% slower for the
# whole process.  A missing library is reported at the first op call.
try:
    load()
except (RuntimeError, OSError):
    pass
