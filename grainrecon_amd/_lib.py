"""ctypes binding of libgrscore.so (the C ABI declared in include/grscore.h).

The library is built in-tree by __graft_entry__.build() / grainrecon_amd.build.
There is no CPU fallback: if the library is missing this module raises.
"""
from __future__ import annotations

import ctypes
import os
import re

from .geometry import CGeometry

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgrscore.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "grscore.h")

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dp = ctypes.POINTER(ctypes.c_double)
c_fp = ctypes.POINTER(ctypes.c_float)
c_ip = ctypes.POINTER(ctypes.c_int32)
c_vp = ctypes.c_void_p

GRS_OK = 0
GRS_ERR_ARG = -1
GRS_ERR_DEVICE = -2
GRS_ERR_STATE = -3
GRS_ERR_NOMEM = -4
METHOD_FAST = 0
METHOD_PLAIN = 1
GT_MATCH = 0
GT_MATCH_FUN = 1
MASK_BITS = 0
MASK_KEYS = 1
MASK_TILED = 2
TUNE_MAX = 24

_SIGS = {
    "grs_version": (ctypes.c_char_p, []),
    "grs_last_error": (ctypes.c_char_p, [c_vp]),
    "grs_create": (c_i32, [c_i32, ctypes.POINTER(c_vp)]),
    "grs_destroy": (None, [c_vp]),
    "grs_set_geometry": (c_i32, [c_vp, ctypes.POINTER(CGeometry), c_dp, c_i64, c_dp, c_i64]),
    "grs_set_distance_stack": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_i64, c_i32]),
    "grs_forward_comp": (c_i32, [c_vp, c_dp, c_dp, c_i64, c_ip, c_ip, c_dp, c_fp, c_i32]),
    "grs_forward_comp_device": (c_i32, [c_vp, c_dp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_i32]),
    "grs_forward_comp_device_noalloc": (c_i32, [c_vp, c_dp, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "grs_grow_indexed_region": (c_i32, [c_vp, c_dp, c_i64, c_dp, c_dp, c_dp, c_i32]),
    "grs_forward_comp_compete": (c_i32, [c_vp, c_dp, c_i64, c_dp, c_i64, c_dp, c_i64, c_dp, c_dp, c_i32]),
    "grs_compete_csr": (c_i32, [c_vp, c_dp, c_i64, ctypes.POINTER(c_i64), c_dp, c_i64, c_i32, ctypes.c_double,
                                ctypes.POINTER(c_i64), c_dp, c_ip, c_ip, c_dp, c_fp]),
    "grs_set_spots": (c_i32, [c_vp, c_i64, c_dp, c_dp, c_dp, c_ip, c_i64]),
    "grs_spot_residual": (c_i32, [c_vp, c_dp, c_dp, c_dp, c_dp, c_dp, c_dp, c_dp]),
    "grs_spot_normal": (c_i32, [c_vp, c_dp, c_dp, ctypes.c_double, c_dp, c_dp, c_dp, c_dp, c_dp, c_dp]),
    "grs_cuda_forward_comp": (c_i32, [c_fp, c_fp, c_i64, c_fp, c_fp, c_fp, c_i64, c_fp, c_i64, c_fp, c_i64, c_fp,
                                      c_fp, c_i64, c_fp]),
    "grs_cuda_grow_indexed_region": (c_i32, [c_fp, c_i64, c_fp, c_fp, c_fp, c_fp, c_i64, c_fp, c_i64, c_fp, c_i64,
                                             c_fp, c_fp, c_i64, c_fp, c_i64, c_fp]),
    "grs_cuda_forward_comp_compete": (c_i32, [c_fp, c_i64, c_fp, c_i64, c_fp, c_i64, c_fp, c_fp, c_fp, c_i64,
                                              c_fp, c_i64, c_fp, c_i64, c_fp, c_fp, c_i64, c_fp, c_i64, c_fp]),
    "grs_render_spots": (c_i32, [c_vp, c_dp, c_ip, c_i64, c_dp, c_i64, c_i32]),
    "grs_simulate_spots": (c_i32, [c_vp, c_dp, c_ip, c_i64, c_dp, c_i64, c_i64, c_ip, c_ip, c_ip, c_dp,
                                   ctypes.POINTER(c_i64)]),
    "grs_set_binary_stack": (c_i32, [c_vp, c_vp, c_i64, c_i64, c_i64, c_i32]),
    "grs_distance_from_mask": (c_i32, [c_vp]),
    "grs_get_mask": (c_i32, [c_vp, c_vp, c_i64]),
    "grs_get_distance_stack": (c_i32, [c_vp, c_fp, c_i64]),
    "grs_device_alloc": (c_i32, [c_vp, c_i64, ctypes.POINTER(c_vp)]),
    "grs_device_free": (c_i32, [c_vp, c_vp]),
    "grs_copy_to_device": (c_i32, [c_vp, c_vp, c_vp, c_i64]),
    "grs_copy_to_host": (c_i32, [c_vp, c_vp, c_vp, c_i64]),
    "grs_device_copy": (c_i32, [c_vp, c_vp, c_vp, c_i64]),
    "grs_set_timing": (c_i32, [c_vp, c_i32]),
    "grs_kernel_time": (c_i32, [c_vp, c_dp, ctypes.POINTER(c_i64), c_i32]),
    "grs_set_audit": (c_i32, [c_vp, c_i32]),
    "grs_set_graphs": (c_i32, [c_vp, c_i32]),
    "grs_set_autotune": (c_i32, [c_vp, c_i32]),
    "grs_get_tuning": (c_i32, [c_vp, c_i32, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32), ctypes.POINTER(c_i32),
                               ctypes.POINTER(c_i32), c_dp]),
    "grs_set_launch_config": (c_i32, [c_vp, c_i32, c_i32, c_i32, c_i32]),
    "grs_graph_stats": (c_i32, [c_vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "grs_fit_or_local": (c_i32, [c_vp, c_dp, c_dp, c_i64, ctypes.c_double, c_i32, c_i32, ctypes.c_double,
                                 ctypes.c_double, c_dp, c_dp, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    "grs_local_sampling": (c_i32, [c_vp, c_dp, c_i64, c_i32, ctypes.c_double, c_dp, c_i64, ctypes.POINTER(c_i64)]),
    "grs_set_orientation_set": (c_i32, [c_vp, c_dp, c_i64]),
    "grs_index_seeds": (c_i32, [c_vp, c_dp, c_i64, c_i32, c_i32, ctypes.c_double, c_i32, ctypes.c_double, c_i32,
                                c_i32, ctypes.c_double, ctypes.c_double, c_dp, ctypes.POINTER(c_i32), c_dp,
                                ctypes.POINTER(c_i32), c_dp]),
    "grs_fit_orientations": (c_i32, [c_dp, c_i64, c_dp, c_dp, ctypes.POINTER(c_i32)]),
    "grs_set_frame_spots": (c_i32, [c_vp, ctypes.POINTER(c_i64), c_i64, c_dp, c_i64]),
    "grs_set_volume": (c_i32, [c_vp, c_i32, c_dp, c_i64, c_i64, c_i64]),
    "grs_forward_gt": (c_i32, [c_vp, c_dp, c_dp, c_i64, c_dp, c_dp, c_i32, c_i32, c_dp]),
    "grs_fit_grain_com": (c_i32, [c_vp, c_dp, ctypes.POINTER(c_i64), c_i64, c_i32, c_i32, ctypes.c_double,
                                  ctypes.c_double, c_dp, c_dp, c_ip, c_ip]),
    "grs_u2euler": (c_i32, [c_dp, c_i64, c_dp]),
    "grs_euler_to_rod_quat": (c_i32, [c_dp, c_i64, c_dp, c_ip]),
    "grs_euler2u_deg": (c_i32, [c_dp, c_i64, c_dp]),
    "grs_get_volume": (c_i32, [c_vp, c_i32, c_dp, c_i64]),
    "grs_get_volume_at": (c_i32, [c_vp, c_i32, ctypes.POINTER(c_i64), c_i64, c_dp]),
    "grs_set_volume_at": (c_i32, [c_vp, c_i32, ctypes.POINTER(c_i64), c_i64, c_dp]),
    "grs_grow_region": (c_i32, [c_vp, c_ip, ctypes.c_double, c_dp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double, c_dp, c_dp, c_ip, ctypes.POINTER(c_i64), c_ip,
                                c_i64, ctypes.POINTER(c_i64)]),
    "grs_pair_spots": (c_i32, [c_vp, c_dp, c_dp, c_i64, c_ip, c_ip, c_fp, ctypes.POINTER(c_i64), c_dp, c_i64]),
    "grs_find_neighbor_ors": (c_i32, [c_vp, ctypes.POINTER(c_i32), c_i64, c_i64, c_i64, ctypes.POINTER(c_i32), c_i64,
                                      c_i32, c_i32, ctypes.POINTER(c_i32), ctypes.POINTER(c_i64),
                                      ctypes.POINTER(c_i64)]),
    "grs_get_audit": (c_i32, [c_vp, ctypes.POINTER(c_i64), c_dp]),
    "grs_math_device": (c_i32, [c_vp, c_i32, c_dp, c_dp, c_i64]),
    "grs_synchronize": (c_i32, [c_vp]),
}

_lib = None


class GrsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"grscore error {code}: {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Loads libgrscore.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m grainrecon_amd.build` "
                               "(the engine has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def header_functions() -> list[str]:
    """Names of the functions declared in include/grscore.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(grs_[a-z0-9_]+)\s*\(", txt)))


def check(code: int, ctx=None) -> None:
    if code != GRS_OK:
        msg = lib().grs_last_error(ctx)
        raise GrsError(code, msg.decode() if msg else "")
