"""ctypes binding of the C ABI declared in ``include/nanodrr_amd.h``.

The library is built in-tree (``nanodrr_amd/libnanodrr_amd.so``). There is no
fallback: when it is missing, every entry point raises.
"""

from __future__ import annotations

import ctypes
import functools
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libnanodrr_amd.so")

NDRR_F32, NDRR_BF16, NDRR_F16 = 0, 1, 2
LAYOUT_LINEAR, LAYOUT_STENCIL, LAYOUT_QUAD, LAYOUT_COEF, LAYOUT_LCOEF, LAYOUT_QCOEF = 0, 1, 2, 3, 4, 5
LAYOUTS = {"linear": LAYOUT_LINEAR, "stencil": LAYOUT_STENCIL, "quad": LAYOUT_QUAD, "coef": LAYOUT_COEF,
           "lcoef": LAYOUT_LCOEF, "qcoef": LAYOUT_QCOEF}
LCOEF_NEGATIVE, LCOEF_CLASS, LCOEF_RANGE = 1, 2, 4  # header word 4 of labelled coefficient records
MAX_CLASSES = 8

_c_int, _c_i64, _c_p, _c_size = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_size_t

# name -> (restype, argtypes); mirrors include/nanodrr_amd.h
SIGNATURES: dict[str, tuple] = {
    "ndrr_version": (_c_int, []),
    "ndrr_strerror": (ctypes.c_char_p, [_c_int]),
    "ndrr_raymarch_fwd": (
        _c_int,
        [_c_p, _c_int, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
         _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_raymarch_fwd_jac": (
        _c_int,
        [_c_p, _c_int, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_raymarch_bwd_jac": (
        _c_int,
        [_c_p, _c_i64, _c_p, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    # the same with (world_to_voxel, rt_inv) in place of M (and the rt_inv gradient in place of gM)
    "ndrr_raymarch_fwd_jac_pose": (
        _c_int,
        [_c_p, _c_int, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_raymarch_bwd_jac_pose": (
        _c_int,
        [_c_p, _c_i64, _c_p, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    # ndrr_raymarch_bwd_jac_pose without its workspace argument (fixed device scratch array)
    "ndrr_raymarch_bwd_jac_pose_noalloc": (
        _c_int,
        [_c_p, _c_i64, _c_p, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_raymarch_bwd_jac_se3": (
        _c_int,
        [_c_p, _c_i64, _c_p, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, ctypes.c_float,
         _c_p, _c_p, _c_int, _c_int, _c_int, _c_int, _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_bwd_workspace_floats": (_c_size, [_c_int, _c_int, _c_int]),
    "ndrr_launch_tile": (None, [ctypes.POINTER(_c_int), ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    "ndrr_raymarch_bwd": (
        _c_int,
        [_c_p, _c_p, _c_int, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p,
         _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_gvol_brick_floats": (_c_size, [_c_int, _c_int, _c_int]),
    "ndrr_raymarch_bwd_bricked": (
        _c_int,
        [_c_p, _c_p, _c_int, _c_int, _c_p, _c_p, _c_p, _c_p, _c_p,
         _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_int,
         _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
         _c_i64, _c_i64, _c_i64, _c_i64, _c_p],
    ),
    "ndrr_labels_to_i8": (_c_int, [_c_p, _c_p, _c_i64, _c_int, _c_p]),
    "ndrr_packed_bytes": (_c_size, [_c_int, _c_int, _c_int, _c_int, _c_int]),
    "ndrr_pack_volume": (_c_int, [_c_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_p, _c_p]),
    "ndrr_pack_coef": (_c_int, [_c_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_p, _c_p]),
    "ndrr_pack_coef_labels": (_c_int, [_c_p, _c_int, _c_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_p, _c_p]),
    "ndrr_packed_label_bytes": (_c_size, [_c_int, _c_int, _c_int, _c_int]),
    "ndrr_pack_labels": (_c_int, [_c_p, _c_int, _c_int, _c_int, _c_int, _c_p, _c_p]),
    "ndrr_make_rays": (_c_int, [_c_p, _c_i64, _c_p, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_int, _c_p, _c_p, _c_p,
                                _c_p]),
    "ndrr_camera_to_voxel": (_c_int, [_c_p, _c_p, _c_p, _c_int, _c_p]),
    "ndrr_camera_to_voxel_bwd": (_c_int, [_c_p, _c_p, _c_p, _c_int, _c_p]),
    "ndrr_conv3x3_fwd": (_c_int, [_c_p, _c_p, _c_p, _c_p] + [_c_int] * 7 + [_c_p]),
    "ndrr_conv3x3_transpose_weights": (_c_int, [_c_p, _c_p, _c_int, _c_int, _c_p]),
    "ndrr_conv3x3_bwd_data": (_c_int, [_c_p, _c_p, _c_p, _c_p] + [_c_int] * 7 + [_c_p]),
    "ndrr_conv3x3_bwd_weight_workspace_floats": (_c_size, [_c_int] * 6),
    "ndrr_conv3x3_bwd_weight": (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_p] + [_c_int] * 6 + [_c_p]),
    "ndrr_head_workspace_floats": (_c_size, [_c_int] * 5),
    "ndrr_head_cast_weights": (_c_int, [_c_p, _c_p, _c_p, _c_int, _c_int, _c_p]),
    "ndrr_head_pool": (_c_int, [_c_p, _c_p, _c_p] + [_c_int] * 4 + [_c_p]),
    "ndrr_head_fwd": (_c_int, [_c_p] * 8 + [_c_int] * 4 + [_c_p]),
    "ndrr_head_bwd": (_c_int, [_c_p] * 14 + [_c_int] * 6 + [_c_p]),
    "ndrr_pose_compose": (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_int, _c_p]),
    "ndrr_pose_compose_bwd": (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_p, _c_p, _c_int, _c_p]),
    "ndrr_adam_step": (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_int, ctypes.c_float, ctypes.c_double, ctypes.c_double]
                       + [ctypes.c_float] * 3 + [_c_p]),
    "ndrr_lattice_offsets": (_c_int, [_c_p, ctypes.c_uint32, _c_p, _c_int, _c_p]),
    "ndrr_adam_step_scheduled": (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_int, _c_p, _c_int, _c_int, _c_p,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_float, _c_p]),
    "ndrr_counter_advance": (_c_int, [_c_p, _c_p]),
    "ndrr_adam_step_scheduled_pair": (_c_int, [_c_p] * 8 + [_c_int, _c_p, _c_int, _c_p, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_float, _c_p]),
    "ndrr_pose_compose_lattice": (_c_int, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_p, _c_int, _c_p, ctypes.c_uint32,
                                           _c_p, _c_p]),
    "ndrr_zncc_workspace_floats": (ctypes.c_int64, [_c_int, _c_int]),
    "ndrr_zncc": (_c_int, [_c_p, _c_p, _c_int, _c_int, ctypes.c_float, _c_p, _c_p, _c_p, _c_p]),
    "ndrr_patch_zncc_workspace_floats": (ctypes.c_int64, [_c_int, _c_int, _c_int, _c_int]),
    "ndrr_patch_zncc": (_c_int, [_c_p, _c_p, _c_int, _c_int, _c_int, _c_int, ctypes.c_float, _c_p, _c_p, _c_p, _c_p]),
}


class NativeError(RuntimeError):
    """A non-zero return code from the native library."""


@functools.cache
def lib() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"nanodrr_amd native library not found at {LIB_PATH}; build it with "
            "`python -m nanodrr_amd.build` (hipcc, gfx950). There is no CPU fallback."
        )
    handle = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    return handle


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().ndrr_strerror(rc).decode()
        raise NativeError(f"{what} failed with code {rc}: {msg}")


def launch_tile() -> tuple[int, int, int]:
    """(tile width, tile height, rays per tile) of the kernels' launch."""
    w, h, r = _c_int(), _c_int(), _c_int()
    lib().ndrr_launch_tile(ctypes.byref(w), ctypes.byref(h), ctypes.byref(r))
    return w.value, h.value, r.value
