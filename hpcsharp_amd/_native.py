"""ctypes binding of libhpcsort.so (the C ABI declared in include/hpcsort.h).

There is no CPU fallback: if the library is missing or no HIP device is
present, every call raises.  Device-resident calls take raw device pointers
(for example torch tensor ``data_ptr()``) and a raw HIP stream handle.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhpcsort.so")

HPCS_OK = 0
HPCS_E_INVALID_ARG = -1
HPCS_E_OOM = -2
HPCS_E_HIP = -3
HPCS_E_BUDGET = -4
HPCS_E_UNSUPPORTED = -5
HPCS_E_BUSY = -6

HPCS_KEY_UINT = 0
HPCS_KEY_FLOAT = 1
HPCS_KEY_INT = 2
HPCS_DTYPE_U32, HPCS_DTYPE_U64, HPCS_DTYPE_F32, HPCS_DTYPE_F64 = 0, 1, 2, 3
HPCS_OPT_RANK, HPCS_OPT_TILE_SHAPE, HPCS_OPT_FALLBACK_ROUNDS, HPCS_OPT_LOOKBACK_WINDOW, HPCS_OPT_PASS_ALGO = 1, 2, 3, 4, 5
HPCS_OPT_DIGIT_BYTES = 7
HPCS_OPT_DIGIT_BITS = 8
HPCS_OPT_GRAPHS = 9
HPCS_OPT_SCRATCH_LIMIT = 12
HPCS_OPT_SEG_RANK = 13
HPCS_OPT_SEG_SMALL = 14
HPCS_OPT_SEG_SHAPE = 15
HPCS_OPT_SEG_MSD = 16
HPCS_OP_SORT, HPCS_OP_TOPK = 0, 1
HPCS_SERVE_REALTIME, HPCS_SERVE_ALONE = 0, 1

# Every symbol include/hpcsort.h declares, with its argument types.
_p, _i64, _i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32


class PlanInfo(ctypes.Structure):
    _fields_ = [
        ("digit_bits", ctypes.c_int32),
        ("passes", ctypes.c_int32),
        ("tile_keys", ctypes.c_int32),
        ("block_threads", ctypes.c_int32),
        ("tiles", ctypes.c_int64),
        ("ideal_bytes", ctypes.c_int64),
        ("scratch_bytes", ctypes.c_int64),
    ]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64), ("total_ms", ctypes.c_double)]


SIGNATURES = {
    "hpcs_version": ([], ctypes.c_int32),
    "hpcs_last_error": ([], ctypes.c_char_p),
    "hpcs_ctx_create": ([_i32, ctypes.POINTER(_p)], _i32),
    "hpcs_ctx_destroy": ([_p], _i32),
    "hpcs_ctx_trim": ([_p], _i32),
    "hpcs_plan_query": ([_p, _i32, _i64, _i32, ctypes.POINTER(PlanInfo)], _i32),
    "hpcs_tile_shape_query": ([_i32, _i32, _i32, ctypes.POINTER(_i32), ctypes.POINTER(_i32)], _i32),
    "hpcs_ctx_set_option": ([_p, _i32, _i64], _i32),
    "hpcs_profile_enable": ([_p, _i32], _i32),
    "hpcs_profile_read": ([_p, ctypes.POINTER(KernelStat), _i32, ctypes.POINTER(_i32)], _i32),
    "hpcs_graph_cache_stats": ([_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64), ctypes.POINTER(_i64)], _i32),
    "hpcs_sort_radix_u32": ([_p, _p, _i64], _i32),
    "hpcs_sort_radix_range_u32": ([_p, _p, _i64, _i64], _i32),
    "hpcs_sort_radix_u64": ([_p, _p, _i64], _i32),
    "hpcs_sort_radix_f32": ([_p, _p, _i64], _i32),
    "hpcs_sort_radix_f64": ([_p, _p, _i64], _i32),
    "hpcs_sort_radix_pairs_u64_u32": ([_p, _p, _p, _i64], _i32),
    "hpcs_sort_radix_pairs_u32_u32": ([_p, _p, _p, _i64], _i32),
    "hpcs_sort_radix_records": ([_p, _p, _i64, _i32, _i32, _i32, _i32], _i32),
    "hpcs_histogram_byte_components_u32": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_histogram_byte_components_u64": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_sort_segments_u32": ([_p, _p, _p, _i32], _i32),
    "hpcs_dev_sort_segments_u32": ([_p, _p, _i64, _p, _i32, _p], _i32),
    "hpcs_dev_sort_radix_u32": ([_p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_u32_noalloc": ([_p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_u64": ([_p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_f32": ([_p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_f64": ([_p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_pairs_u64_u32": ([_p, _p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_pairs_u32_u32": ([_p, _p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_records": ([_p, _p, _i64, _i32, _i32, _i32, _i32, _p], _i32),
    "hpcs_dev_histogram_byte_components_u32": ([_p, _p, _i64, _p, _p], _i32),
    "hpcs_dev_histogram_byte_components_u64": ([_p, _p, _i64, _p, _p], _i32),
    "hpcs_ctx_reserve": ([_p, _i32, _i64], _i32),
    "hpcs_ctx_reserve_segments": ([_p, _i64, _i32], _i32),
    "hpcs_ctx_load_tuning": ([_p, ctypes.c_char_p], _i32),
    "hpcs_tuning_check": ([ctypes.c_char_p, _p], _i32),
    "hpcs_serve_run": ([_i32, _i32, _i64, _p, _i64, _i32, _p, _p], _i32),
    "hpcs_serve_run_config": ([_i32, _p, _p, _i64, _p, _p], _i32),
    "hpcs_serve_last_error": ([], ctypes.c_char_p),
    "hpcs_sort_radix_keys_items": ([_p, _p, _i32, _p, _i32, _i64], _i32),
    "hpcs_dev_sort_radix_keys_items": ([_p, _p, _i32, _p, _i32, _i64, _p], _i32),
    "hpcs_histogram_nbit_components_u32": ([_p, _p, _i64, _i64, _i32, _p], _i32),
    "hpcs_histogram_nbit_components_u64": ([_p, _p, _i64, _i64, _i32, _p], _i32),
    "hpcs_dev_histogram_nbit_components_u32": ([_p, _p, _i64, _i32, _p, _p], _i32),
    "hpcs_dev_histogram_nbit_components_u64": ([_p, _p, _i64, _i32, _p, _p], _i32),
    "hpcs_histogram_byte_components_records": ([_p, _p, _i64, _i64, _i32, _i32, _i32, _p], _i32),
    "hpcs_dev_histogram_byte_components_records": ([_p, _p, _i64, _i32, _i32, _i32, _p, _p], _i32),
    "hpcs_sort_radix_i32": ([_p, _p, _i64], _i32),
    "hpcs_sort_radix_i64": ([_p, _p, _i64], _i32),
    "hpcs_dev_sort_radix_i32": ([_p, _p, _i64, _p], _i32),
    "hpcs_dev_sort_radix_i64": ([_p, _p, _i64, _p], _i32),
    "hpcs_histogram_byte_components_i32": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_histogram_byte_components_i64": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_dev_histogram_byte_components_i32": ([_p, _p, _i64, _p, _p], _i32),
    "hpcs_dev_histogram_byte_components_i64": ([_p, _p, _i64, _p, _p], _i32),
    "hpcs_topk_i32": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_topk_i64": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_dev_topk_i32": ([_p, _p, _i64, _i64, _p, _p], _i32),
    "hpcs_dev_topk_i64": ([_p, _p, _i64, _i64, _p, _p], _i32),
    "hpcs_select_topk_u32": ([_p, _p, _i64, _i64, _i64], _i32),
    "hpcs_topk_u32": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_topk_u64": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_topk_f32": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_topk_f64": ([_p, _p, _i64, _i64, _p], _i32),
    "hpcs_dev_topk_u32": ([_p, _p, _i64, _i64, _p, _p], _i32),
    "hpcs_dev_topk_u64": ([_p, _p, _i64, _i64, _p, _p], _i32),
    "hpcs_dev_topk_f32": ([_p, _p, _i64, _i64, _p, _p], _i32),
    "hpcs_dev_topk_f64": ([_p, _p, _i64, _i64, _p, _p], _i32),
}

_lib = None
_lock = threading.Lock()
_contexts: dict[int, ctypes.c_void_p] = {}


class HpcsError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{message} (status {code})")
        self.code = code


# The tuning table measured on MI355X by hpcsharp_amd.autotune; the
# process-wide contexts load it unless HPCS_TUNE_FILE names another.
TUNE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_gfx950.txt")


def lib() -> ctypes.CDLL:
    """Load libhpcsort.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m hpcsharp_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != HPCS_OK:
        msg = lib().hpcs_last_error()
        raise HpcsError(rc, msg.decode() if msg else "hpcsort error")


def context(device: int = 0) -> ctypes.c_void_p:
    """The process-wide context of a device, created on first use."""
    with _lock:
        ctx = _contexts.get(device)
        if ctx is None:
            ctx = ctypes.c_void_p()
            check(lib().hpcs_ctx_create(device, ctypes.byref(ctx)))
            if not os.environ.get("HPCS_TUNE_FILE") and os.path.exists(TUNE_PATH):
                check(lib().hpcs_ctx_load_tuning(ctx, TUNE_PATH.encode()))
            _contexts[device] = ctx
        return ctx


def load_tuning(path: str | None, ctx=None) -> None:
    """Replace a context's tuning table (None clears it)."""
    check(lib().hpcs_ctx_load_tuning(ctx if ctx is not None else context(), (path or "").encode()))


def new_context(device: int = 0) -> ctypes.c_void_p:
    """A private context (its own scratch and stream), e.g. one per serving stream."""
    ctx = ctypes.c_void_p()
    check(lib().hpcs_ctx_create(device, ctypes.byref(ctx)))
    return ctx


def destroy_context(ctx: ctypes.c_void_p) -> None:
    check(lib().hpcs_ctx_destroy(ctx))


def plan_query(dtype: int, n: int, with_values: bool = False, device: int = 0) -> PlanInfo:
    info = PlanInfo()
    check(lib().hpcs_plan_query(context(device), dtype, n, 1 if with_values else 0, ctypes.byref(info)))
    return info


def profile_enable(on: bool = True, ctx=None) -> None:
    check(lib().hpcs_profile_enable(ctx if ctx is not None else context(), 1 if on else 0))


def profile_read(ctx=None) -> dict[str, tuple[int, float]]:
    """{kernel name: (launches, total ms)} since profiling was enabled."""
    buf = (KernelStat * 64)()
    cnt = ctypes.c_int32()
    check(lib().hpcs_profile_read(ctx if ctx is not None else context(), buf, 64, ctypes.byref(cnt)))
    return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(cnt.value)}


def graph_cache_stats(ctx=None) -> tuple[int, int, int]:
    """(graphs held, captures, graph launches) of the context's graph cache."""
    e, c, r = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().hpcs_graph_cache_stats(ctx if ctx is not None else context(), ctypes.byref(e), ctypes.byref(c),
                                       ctypes.byref(r)))
    return e.value, c.value, r.value


def set_option(option: int, value: int, ctx=None) -> None:
    check(lib().hpcs_ctx_set_option(ctx if ctx is not None else context(), option, value))
