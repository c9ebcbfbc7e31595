"""ctypes binding of ``libstablebear_amd.so`` (the C-ABI in
``include/stablebear_amd.h``).

The library is the only compute path: importing this module fails loudly if
the HIP library has not been built (``python -m stablebear_amd.build``).
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstablebear_amd.so")

OK = 0
ERR_INVALID = 1
ERR_RUNTIME = 2
ERR_NODEVICE = 3
ERR_CANCELLED = 4
TIMEOUT = 5
ERR_NOMEM = 6

F32 = 1
F64 = 2

OP_L1 = 0
OP_LP = 1
OP_DOT = 2

OUT_DISTMAT = 0
OUT_SYMMAT = 1
OUT_DENSE = 2

PREC_FP64 = 0
PREC_FP32C = 1

c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_dblp = ctypes.POINTER(ctypes.c_double)
vp = ctypes.c_void_p


class PcfSet(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("offsets", c_i64p), ("t", vp), ("v", vp),
                ("dtype", ctypes.c_int32)]


HOST = 0
DEVICE = 1


class PcfView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("offsets", vp), ("t", vp), ("v", vp),
                ("t_stride", ctypes.c_int64), ("v_stride", ctypes.c_int64),
                ("dtype", ctypes.c_int32), ("location", ctypes.c_int32)]


IO_TENSOR = 1
IO_OBJECT = 2


class IoInfo(ctypes.Structure):
    _fields_ = [("version", ctypes.c_int32), ("kind", ctypes.c_uint32),
                ("base_format", ctypes.c_int32), ("sub_format", ctypes.c_int32),
                ("ndim", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("shape", ctypes.c_uint64 * 32), ("elements", ctypes.c_uint64),
                ("payload", ctypes.c_uint64)]


class TileShape(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int32), ("cols", ctypes.c_int32), ("cap", ctypes.c_int32),
                ("lane_pairs", ctypes.c_int32)]


class PlanStats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_uint64), ("segments", ctypes.c_uint64),
                ("out_count", ctypes.c_uint64), ("tiles", ctypes.c_uint64),
                ("input_bytes", ctypes.c_uint64), ("rows", ctypes.c_int32),
                ("cols", ctypes.c_int32), ("cap", ctypes.c_int32), ("problems", ctypes.c_int32),
                ("long_pairs", ctypes.c_int32), ("lane_pairs", ctypes.c_int32)]


# name -> (restype, argtypes); restype None means int status
_SIGS = {
    "sbamd_last_error": (ctypes.c_char_p, []),
    "sbamd_version": (ctypes.c_char_p, []),
    "sbamd_device_count": (None, [ctypes.POINTER(ctypes.c_int)]),
    "sbamd_ctx_create": (None, [ctypes.c_int, ctypes.POINTER(vp)]),
    "sbamd_ctx_destroy": (None, [vp]),
    "sbamd_ctx_synchronize": (None, [vp]),
    "sbamd_ctx_stream": (None, [vp, ctypes.POINTER(vp)]),
    "sbamd_set_create": (None, [vp, ctypes.POINTER(PcfSet), ctypes.c_int32, ctypes.POINTER(vp)]),
    "sbamd_set_create_gpu": (None, [vp, ctypes.POINTER(PcfView), ctypes.c_int32,
                                    ctypes.POINTER(vp)]),
    "sbamd_set_export": (None, [vp, vp, c_i64p, c_i64p, c_i64p, c_i32p]),
    "sbamd_rows_reduce": (None, [vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, vp, ctypes.c_int32,
                                 vp, vp, vp, vp]),
    "sbamd_rows_knn": (None, [vp, ctypes.c_int32, ctypes.c_int64, ctypes.c_int64, vp, ctypes.c_int32,
                              ctypes.c_int32, vp, vp, vp]),
    "sbamd_set_create_image": (None, [vp, vp, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(vp)]),
    "sbamd_io_parse": (None, [vp, ctypes.c_uint64, ctypes.POINTER(IoInfo)]),
    "sbamd_io_index": (None, [vp, ctypes.c_uint64, ctypes.POINTER(IoInfo), c_u64p, c_u64p, c_u64p]),
    "sbamd_io_header": (None, [ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                               c_u64p, vp, ctypes.c_uint64, c_u64p]),
    "sbamd_io_write_pcfs": (None, [ctypes.c_int64, c_i64p, vp, vp, ctypes.c_int32, vp,
                                   ctypes.c_uint64, c_u64p]),
    "sbamd_set_destroy": (None, [vp]),
    "sbamd_set_info": (None, [vp, c_i64p, c_i64p, c_i64p]),
    "sbamd_set_update_values": (None, [vp, vp, vp]),
    "sbamd_set_lp_norm": (None, [vp, ctypes.c_double, vp, ctypes.c_int32, vp]),
    "sbamd_plan_create": (None, [vp, vp, vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                 ctypes.POINTER(TileShape), ctypes.POINTER(vp)]),
    "sbamd_plan_create_batched": (None, [vp, ctypes.c_int32, ctypes.POINTER(vp), ctypes.c_int32,
                                         ctypes.c_double, ctypes.c_int32,
                                         ctypes.POINTER(TileShape), ctypes.POINTER(vp)]),
    "sbamd_plan_destroy": (None, [vp]),
    "sbamd_plan_info": (None, [vp, ctypes.POINTER(PlanStats)]),
    "sbamd_plan_out_offset": (None, [vp, ctypes.c_int32, c_u64p]),
    "sbamd_plan_set_fold_threshold": (None, [vp, ctypes.c_int64]),
    "sbamd_plan_set_persistent": (None, [vp, ctypes.c_int32]),
    "sbamd_plan_set_long_pairs": (None, [vp, ctypes.c_int32]),
    "sbamd_plan_execute": (None, [vp, vp, ctypes.c_int32, vp, vp]),
    "sbamd_plan_backward": (None, [vp, vp, vp, ctypes.c_int32, vp, vp, ctypes.c_int32, vp]),
    "sbamd_ctx_set_output_budget": (None, [vp, ctypes.c_uint64]),
    "sbamd_ctx_set_tile_shape": (None, [vp, ctypes.POINTER(TileShape)]),
    "sbamd_plan_backward_noalloc": (None, [vp, vp, vp, ctypes.c_int32, vp, vp, ctypes.c_int32, vp]),
    "sbamd_plan_execute_to_host": (None, [vp, vp, ctypes.c_int32]),
    "sbamd_plain_create": (None, [vp, ctypes.POINTER(PcfSet), ctypes.POINTER(vp)]),
    "sbamd_plain_execute": (None, [vp, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, vp, vp]),
    "sbamd_plain_destroy": (None, [vp]),
    "sbamd_graph_create": (None, [vp, vp, ctypes.c_int32, ctypes.POINTER(vp)]),
    "sbamd_graph_launch": (None, [vp, vp]),
    "sbamd_graph_destroy": (None, [vp]),
    "sbamd_autotune": (None, [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(TileShape),
                              c_dblp]),
    "sbamd_plan_set_shape": (None, [vp, ctypes.POINTER(TileShape)]),
    "sbamd_plan_class_shapes": (None, [vp, ctypes.POINTER(TileShape), c_i64p]),
    "sbamd_autotune_cache_stats": (None, [c_u64p, c_u64p]),
    "sbamd_plan_set_class_shapes": (None, [vp, ctypes.POINTER(TileShape)]),
    "sbamd_pairwise_submit": (None, [vp, ctypes.POINTER(PcfSet), ctypes.POINTER(PcfSet),
                                     ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                     ctypes.c_int32, vp, ctypes.c_int32, ctypes.POINTER(vp)]),
    "sbamd_pairwise_submit_view": (None, [vp, ctypes.POINTER(PcfView), ctypes.POINTER(PcfView),
                                          ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                          ctypes.c_int32, vp, ctypes.c_int32, ctypes.POINTER(vp)]),
    "sbamd_job_wait": (None, [vp, ctypes.c_int64]),
    "sbamd_job_progress": (None, [vp, c_u64p, c_u64p]),
    "sbamd_job_cancel": (None, [vp]),
    "sbamd_job_destroy": (None, [vp]),
    "sbamd_lp_distance": (None, [vp, ctypes.POINTER(PcfSet), ctypes.POINTER(PcfSet),
                                 ctypes.c_double, c_dblp]),
    "sbamd_lp_norm": (None, [vp, ctypes.POINTER(PcfSet), ctypes.c_double, ctypes.c_int32, vp,
                             ctypes.c_int32]),
    "sbamd_iterate_rectangles": (None, [vp, ctypes.POINTER(PcfSet), ctypes.POINTER(PcfSet),
                                        ctypes.c_double, ctypes.c_double, vp, ctypes.c_int64,
                                        c_i64p]),
    "sbamd_validate_set": (None, [ctypes.POINTER(PcfSet)]),
    "sbamd_count_segments": (None, [ctypes.POINTER(PcfSet), ctypes.POINTER(PcfSet),
                                    ctypes.c_int32, c_u64p]),
    "sbamd_plan_tiles_host": (None, [ctypes.c_int64, c_i64p, ctypes.c_int64, c_i64p,
                                     ctypes.c_int32, ctypes.c_int32, c_i32p, ctypes.c_int64,
                                     c_i64p, c_i64p]),
    "sbamd_queue_order_host": (None, [ctypes.c_int64, c_i32p, c_i64p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, c_i64p, c_i64p]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """The loaded library (raises RuntimeError when it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"stablebear_amd HIP library not found at {LIB_PATH}; build it with "
                "`python -m stablebear_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res if res is not None else ctypes.c_int
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().sbamd_last_error()
    return msg.decode("utf-8", "replace") if msg else ""


class CancelledError(RuntimeError):
    """The job was cancelled before it finished."""


def check(rc: int) -> None:
    """Raise the Python exception matching a C-ABI status."""
    if rc == OK:
        return
    msg = last_error()
    if rc == ERR_INVALID:
        raise ValueError(msg)
    if rc == ERR_CANCELLED:
        raise CancelledError(msg or "cancelled")
    if rc == ERR_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg or f"stablebear_amd error {rc}")


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().sbamd_device_count(ctypes.byref(n)))
    return n.value
