"""ctypes binding of libxpbd_hip.so (the C ABI declared in include/xpbd_abi.h).

The library is required: importing the solver without it raises, there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_NAME = "libxpbd_hip.so"
# XPBD_LIB selects another build of the library (for example a debug build).
LIB_PATH = os.environ.get("XPBD_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                      LIB_NAME)

XPBD_OK = 0
XPBD_EINVAL = -1
XPBD_EHIP = -2
XPBD_ENOMEM = -3
XPBD_ESTATE = -4
XPBD_ENUMERIC = -5
XPBD_MAX_SDF = 8
XPBD_ABI_VERSION = 5

ACCUM_ATOMIC = 0
ACCUM_CSR = 1
SDF_PLANE = 0
SDF_SPHERE = 1
SDF_CUBE = 2
SDF_HANG = 3
MAX_HANG_POINTS = 4096
PATH_KERNELS = 0
PATH_GRAPH = 1
PATH_PERSISTENT = 2
PATH_AUTO = 3
LAYOUT_FLAT = 0
LAYOUT_PATCH = 1
ORDER_NONE = 0
ORDER_BISECT = 1
ORDER_MORTON = 2
NOALLOC_MAX_SCENES = 64
NOALLOC_LIST_BYTES = 1 << 20

# Every symbol the header declares; tests check that the library exports them.
EXPORTED = (
    "xpbd_abi_info", "xpbd_desc_init", "xpbd_create", "xpbd_destroy", "xpbd_update_params",
    "xpbd_set_sdfs", "xpbd_set_gravity", "xpbd_compute_lumped_mass",
    "xpbd_get_positions", "xpbd_set_positions", "xpbd_get_positions_device",
    "xpbd_set_positions_device", "xpbd_get_velocities", "xpbd_set_velocities",
    "xpbd_get_predicted", "xpbd_get_masses", "xpbd_get_degenerate_tets", "xpbd_get_multipliers",
    "xpbd_get_visual_indices", "xpbd_apply_external_forces", "xpbd_solve_sdf_collision",
    "xpbd_solve_step", "xpbd_advance", "xpbd_substep", "xpbd_frame", "xpbd_synchronize",
    "xpbd_get_stream", "xpbd_set_logging", "xpbd_log_size", "xpbd_get_log", "xpbd_clear_log",
    "xpbd_get_state", "xpbd_set_state", "xpbd_get_sizes", "xpbd_set_exec_path",
    "xpbd_set_kernel_timing", "xpbd_get_kernel_time", "xpbd_check_finite", "xpbd_last_error",
    "xpbd_patch_stats", "xpbd_get_device_bytes", "xpbd_batch_opts_init", "xpbd_batch_create",
    "xpbd_batch_destroy", "xpbd_batch_set_members", "xpbd_batch_set_members_noalloc",
    "xpbd_batch_substep", "xpbd_batch_frame",
    "xpbd_batch_synchronize", "xpbd_batch_get_stream", "xpbd_batch_info", "xpbd_batch_grid",
    "xpbd_batch_path", "xpbd_set_barrier_timeout", "xpbd_batch_set_barrier_timeout",
)


class XpbdSdf(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("fixed", ctypes.c_int32),
        ("p", ctypes.c_double * 4),
        ("dir", ctypes.c_double * 3),
        ("points", ctypes.POINTER(ctypes.c_float)),
        ("n_points", ctypes.c_int32),
        ("pad", ctypes.c_int32),
        ("vel", ctypes.c_double * 3),
    ]


class XpbdDesc(ctypes.Structure):
    _fields_ = [
        ("n_verts", ctypes.c_int32),
        ("n_tets", ctypes.c_int32),
        ("rest_positions", ctypes.POINTER(ctypes.c_float)),
        ("tets", ctypes.POINTER(ctypes.c_int32)),
        ("young_modulus", ctypes.c_double),
        ("poisson_ratio", ctypes.c_double),
        ("scale", ctypes.c_double),
        ("offset", ctypes.c_double * 3),
        ("frame_dt", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("substeps_num", ctypes.c_int32),
        ("constraints_iter", ctypes.c_int32),
        ("hydrostatic_relaxation", ctypes.c_double),
        ("deviatoric_relaxation", ctypes.c_double),
        ("cheb_enable", ctypes.c_int32),
        ("cheb_warmup", ctypes.c_int32),
        ("cheb_gamma", ctypes.c_double),
        ("block_neohookean_enable", ctypes.c_int32),
        ("monitor_convergence", ctypes.c_int32),
        ("damping", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("total_mass", ctypes.c_double),
        ("n_sdf", ctypes.c_int32),
        ("sdfs", ctypes.POINTER(XpbdSdf)),
        ("accumulate_mode", ctypes.c_int32),
        ("reorder", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("layout", ctypes.c_int32),
    ]


class XpbdBatchOpts(ctypes.Structure):
    _fields_ = [
        ("graph", ctypes.c_int32),
        ("graph_substeps", ctypes.c_int32),
        ("vertex_block", ctypes.c_int32),
        ("grid_pct", ctypes.c_int32),
        ("xcd_split", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
    ]


class XpbdError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"xpbd error {code}: {message}")
        self.code = code


_lib = None


def library_available() -> bool:
    return os.path.exists(LIB_PATH)


def load() -> ctypes.CDLL:
    """Loads libxpbd_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if os.environ.get("XPBD_NO_TORCH") != "1":
        # PyTorch-ROCm ships its own HIP runtime (SONAME libamdhip64.so.7).  When
        # it is loaded first, libxpbd_hip.so binds to that same runtime, so torch
        # tensors and this library share one device context; loaded the other way
        # round the process would hold two HIP runtimes and torch would see no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python rdc-xpbd-taichi_amd/build.py` "
            "(the solver has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    pp = ctypes.POINTER(ctypes.c_void_p)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    i32 = ctypes.c_int32
    sigs = {
        "xpbd_abi_info": [ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_size_t),
                          ctypes.POINTER(ctypes.c_size_t)],
        "xpbd_desc_init": [ctypes.POINTER(XpbdDesc)],
        "xpbd_create": [ctypes.POINTER(XpbdDesc), pp],
        "xpbd_destroy": [vp],
        "xpbd_update_params": [vp, ctypes.POINTER(XpbdDesc)],
        "xpbd_set_sdfs": [vp, ctypes.POINTER(XpbdSdf), i32],
        "xpbd_set_gravity": [vp, dp],
        "xpbd_compute_lumped_mass": [vp, ctypes.c_double],
        "xpbd_get_positions": [vp, fp],
        "xpbd_set_positions": [vp, fp],
        "xpbd_get_positions_device": [vp, vp],
        "xpbd_set_positions_device": [vp, vp],
        "xpbd_get_velocities": [vp, fp],
        "xpbd_set_velocities": [vp, fp],
        "xpbd_get_predicted": [vp, fp],
        "xpbd_get_masses": [vp, fp, fp],
        "xpbd_get_multipliers": [vp, fp, fp],
        "xpbd_get_degenerate_tets": [vp, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32)],
        "xpbd_get_visual_indices": [vp, ctypes.POINTER(ctypes.c_uint32)],
        "xpbd_apply_external_forces": [vp, ctypes.c_double],
        "xpbd_solve_sdf_collision": [vp, ctypes.c_double],
        "xpbd_solve_step": [vp, ctypes.c_double],
        "xpbd_advance": [vp, ctypes.c_double, ctypes.c_double],
        "xpbd_substep": [vp, ctypes.c_double],
        "xpbd_frame": [vp, i32],
        "xpbd_synchronize": [vp],
        "xpbd_get_stream": [vp, pp],
        "xpbd_set_logging": [vp, i32],
        "xpbd_log_size": [vp, ctypes.POINTER(ctypes.c_int64)],
        "xpbd_get_log": [vp, dp, ctypes.c_int64],
        "xpbd_clear_log": [vp],
        "xpbd_get_state": [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                           dp, dp],
        "xpbd_set_state": [vp, i32, ctypes.c_double, ctypes.c_double],
        "xpbd_get_sizes": [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)],
        "xpbd_set_exec_path": [vp, i32],
        "xpbd_set_kernel_timing": [vp, i32],
        "xpbd_get_kernel_time": [vp, dp, ctypes.POINTER(ctypes.c_int64)],
        "xpbd_check_finite": [vp],
        "xpbd_last_error": [ctypes.c_char_p, ctypes.c_size_t],
        "xpbd_get_device_bytes": [vp, ctypes.POINTER(ctypes.c_int64)],
        "xpbd_batch_opts_init": [ctypes.POINTER(XpbdBatchOpts)],
        "xpbd_batch_create": [ctypes.POINTER(ctypes.c_void_p), i32, ctypes.POINTER(XpbdBatchOpts),
                              pp],
        "xpbd_batch_destroy": [vp],
        "xpbd_batch_set_members": [vp, ctypes.POINTER(ctypes.c_void_p), i32],
        "xpbd_batch_set_members_noalloc": [vp, ctypes.POINTER(ctypes.c_void_p), i32],
        "xpbd_batch_substep": [vp],
        "xpbd_batch_frame": [vp, i32],
        "xpbd_batch_synchronize": [vp],
        "xpbd_batch_get_stream": [vp, pp],
        "xpbd_batch_info": [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)],
        "xpbd_batch_grid": [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                            ctypes.POINTER(ctypes.c_int32)],
        "xpbd_batch_path": [vp, ctypes.POINTER(ctypes.c_int32)],
        "xpbd_set_barrier_timeout": [vp, ctypes.c_int64],
        "xpbd_batch_set_barrier_timeout": [vp, ctypes.c_int64],
        "xpbd_patch_stats": [fp, ctypes.POINTER(ctypes.c_int32), i32, i32,
                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int64),
                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)],
    }
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    ver = ctypes.c_int32()
    dsz = ctypes.c_size_t()
    ssz = ctypes.c_size_t()
    lib.xpbd_abi_info(ctypes.byref(ver), ctypes.byref(dsz), ctypes.byref(ssz))
    if (ver.value != XPBD_ABI_VERSION or dsz.value != ctypes.sizeof(XpbdDesc)
            or ssz.value != ctypes.sizeof(XpbdSdf)):
        raise ImportError(
            f"{LIB_PATH}: ABI mismatch (version {ver.value}, desc {dsz.value} B vs "
            f"{ctypes.sizeof(XpbdDesc)} B, sdf {ssz.value} B vs {ctypes.sizeof(XpbdSdf)} B)")
    _lib = lib
    return lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    load().xpbd_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc: int):
    if rc != XPBD_OK:
        msg = last_error()
        if rc == XPBD_EINVAL:
            raise ValueError(msg)
        raise XpbdError(rc, msg)


def fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def patch_stats(tets: np.ndarray, n_verts: int, positions=None) -> dict:
    """Patch layout the library would build for these tets (host only).  With
    positions, the solver's bisection ordering is applied first."""
    lib = load()
    t = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    pos = None if positions is None else np.ascontiguousarray(positions, dtype=np.float32)
    npch = ctypes.c_int32()
    nsl = ctypes.c_int64()
    mv = ctypes.c_int32()
    mt = ctypes.c_int32()
    check(lib.xpbd_patch_stats(None if pos is None else fptr(pos),
                               t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), t.shape[0],
                               int(n_verts), ctypes.byref(npch), ctypes.byref(nsl),
                               ctypes.byref(mv), ctypes.byref(mt)))
    return {"n_patches": npch.value, "n_slots": nsl.value, "max_patch_verts": mv.value,
            "max_patch_tets": mt.value}
