"""ctypes loader for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path
(hpcsharp_amd/libhpcsort.so) never imports or calls anything here.

Every wrapper works on numpy arrays in place, the way the reference's
extension methods work on the caller's managed array.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

KEY_UINT = 0
KEY_FLOAT = 1
KEY_INT = 2


def build() -> str:
    """Compile oracle/hpcs_oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        p, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        sigs = {
            "hpcs_oracle_histogram_byte_components_u32": (p, i64, i64, p),
            "hpcs_oracle_histogram_byte_components_u64": (p, i64, i64, p),
            "hpcs_oracle_histogram_nbit_components_u32": (p, i64, i64, ctypes.c_int, p),
            "hpcs_oracle_histogram_nbit_components_u64": (p, i64, i64, ctypes.c_int, p),
            "hpcs_oracle_histogram_byte_components_par_u32": (p, i64, i64, i64, p),
            "hpcs_oracle_histogram_byte_components_par_u64": (p, i64, i64, i64, p),
            "hpcs_oracle_sort_radix_lsd_u32": (p, i64, i64),
            "hpcs_oracle_sort_radix_lsd_u64": (p, i64, i64),
            "hpcs_oracle_sort_radix_lsd_nbit_u32": (p, i64, i64, ctypes.c_int),
            "hpcs_oracle_sort_radix_lsd_nbit_u64": (p, i64, i64, ctypes.c_int),
            "hpcs_oracle_sort_radix_par_u32": (p, i64, i64),
            "hpcs_oracle_sort_radix_par_u64": (p, i64),
            "hpcs_oracle_sort_radix_f32": (p, i64),
            "hpcs_oracle_sort_radix_f64": (p, i64),
            "hpcs_oracle_sort_radix_records": (p, i64, i32, i32, i32, i32),
            "hpcs_oracle_sort_radix_pairs_u64_u32": (p, p, i64),
            "hpcs_oracle_sort_radix_pairs_u32_u32": (p, p, i64),
            "hpcs_oracle_sort_segments_u32": (p, p, i32),
            "hpcs_oracle_select_topk_u32": (p, i64, i64),
            "hpcs_oracle_histogram_byte_components_i32": (p, i64, i64, p),
            "hpcs_oracle_histogram_byte_components_i64": (p, i64, i64, p),
            "hpcs_oracle_sort_radix_lsd_i32": (p, i64),
            "hpcs_oracle_sort_radix_lsd_i64": (p, i64),
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = list(args)
            f.restype = ctypes.c_int
        for name, t in (("hpcs_oracle_f32_key", ctypes.c_uint32), ("hpcs_oracle_f32_unkey", ctypes.c_uint32),
                        ("hpcs_oracle_f64_key", ctypes.c_uint64), ("hpcs_oracle_f64_unkey", ctypes.c_uint64)):
            f = getattr(L, name)
            f.argtypes = [t]
            f.restype = t
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def histogram_byte_components_u32(a: np.ndarray, l: int, r: int) -> np.ndarray:
    out = np.zeros((4, 256), dtype=np.int32)
    _check(lib().hpcs_oracle_histogram_byte_components_u32(_ptr(a), l, r, _ptr(out)), "histogram u32")
    return out


def histogram_byte_components_u64(a: np.ndarray, l: int, r: int) -> np.ndarray:
    out = np.zeros((8, 256), dtype=np.int32)
    _check(lib().hpcs_oracle_histogram_byte_components_u64(_ptr(a), l, r, _ptr(out)), "histogram u64")
    return out


def histogram_nbit_components(a: np.ndarray, l: int, r: int, bits: int) -> np.ndarray:
    width = 32 if a.dtype == np.uint32 else 64
    ndig = (width + bits - 1) // bits
    out = np.zeros((ndig, 1 << bits), dtype=np.int32)
    f = lib().hpcs_oracle_histogram_nbit_components_u32 if width == 32 else lib().hpcs_oracle_histogram_nbit_components_u64
    _check(f(_ptr(a), l, r, bits, _ptr(out)), "histogram nbit")
    return out


def histogram_byte_components_par(a: np.ndarray, l: int, r: int, threshold: int = 16 * 1024) -> np.ndarray:
    if a.dtype == np.uint32:
        out = np.zeros((4, 256), dtype=np.uint32)
        _check(lib().hpcs_oracle_histogram_byte_components_par_u32(_ptr(a), l, r, threshold, _ptr(out)), "histogram par")
    else:
        out = np.zeros((8, 256), dtype=np.uint32)
        _check(lib().hpcs_oracle_histogram_byte_components_par_u64(_ptr(a), l, r, threshold, _ptr(out)), "histogram par")
    return out


def sort_radix_lsd(a: np.ndarray, start: int = 0, length: int | None = None) -> np.ndarray:
    if length is None:
        length = a.size - start
    f = lib().hpcs_oracle_sort_radix_lsd_u32 if a.dtype == np.uint32 else lib().hpcs_oracle_sort_radix_lsd_u64
    _check(f(_ptr(a), start, length), "sort lsd")
    return a


def sort_radix_lsd_nbit(a: np.ndarray, bits: int, start: int = 0, length: int | None = None) -> np.ndarray:
    if length is None:
        length = a.size - start
    f = lib().hpcs_oracle_sort_radix_lsd_nbit_u32 if a.dtype == np.uint32 else lib().hpcs_oracle_sort_radix_lsd_nbit_u64
    _check(f(_ptr(a), start, length, bits), "sort nbit")
    return a


def sort_radix_par(a: np.ndarray, quantum: int = 64 * 1024) -> np.ndarray:
    if a.dtype == np.uint32:
        _check(lib().hpcs_oracle_sort_radix_par_u32(_ptr(a), a.size, quantum), "sort par u32")
    else:
        _check(lib().hpcs_oracle_sort_radix_par_u64(_ptr(a), a.size), "sort par u64")
    return a


def sort_radix_float(a: np.ndarray) -> np.ndarray:
    if a.dtype == np.float32:
        _check(lib().hpcs_oracle_sort_radix_f32(_ptr(a), a.size), "sort f32")
    else:
        _check(lib().hpcs_oracle_sort_radix_f64(_ptr(a), a.size), "sort f64")
    return a


def sort_radix_records(recs: np.ndarray, key_offset: int, key_bits: int, key_kind: int = KEY_UINT) -> np.ndarray:
    rec_size = recs.dtype.itemsize
    _check(lib().hpcs_oracle_sort_radix_records(_ptr(recs), recs.size, rec_size, key_offset, key_bits, key_kind),
           "sort records")
    return recs


def sort_radix_pairs(keys: np.ndarray, vals: np.ndarray):
    f = lib().hpcs_oracle_sort_radix_pairs_u64_u32 if keys.dtype == np.uint64 else lib().hpcs_oracle_sort_radix_pairs_u32_u32
    _check(f(_ptr(keys), _ptr(vals), keys.size), "sort pairs")
    return keys, vals


def sort_segments_u32(a: np.ndarray, seg_begin: np.ndarray) -> np.ndarray:
    seg_begin = np.ascontiguousarray(seg_begin, dtype=np.int64)
    _check(lib().hpcs_oracle_sort_segments_u32(_ptr(a), _ptr(seg_begin), seg_begin.size - 1), "sort segments")
    return a


def f32_key(bits: int) -> int:
    return lib().hpcs_oracle_f32_key(bits)


def f64_key(bits: int) -> int:
    return lib().hpcs_oracle_f64_key(bits)


def select_topk_u32(a: np.ndarray, k: int) -> np.ndarray:
    """SelectTopK(uint[], k): a[k] becomes the k-th smallest, a[k:] the n-k largest."""
    _check(lib().hpcs_oracle_select_topk_u32(_ptr(a), a.size, k), "select topk")
    return a


def histogram_byte_components_signed(a: np.ndarray, l: int, r: int) -> np.ndarray:
    """HistogramByteComponents(int[] / long[], l, r): the top digit is byte ^ 128."""
    if a.dtype == np.int32:
        out = np.zeros((4, 256), dtype=np.int32)
        _check(lib().hpcs_oracle_histogram_byte_components_i32(_ptr(a), l, r, _ptr(out)), "histogram i32")
    else:
        out = np.zeros((8, 256), dtype=np.int32)
        _check(lib().hpcs_oracle_histogram_byte_components_i64(_ptr(a), l, r, _ptr(out)), "histogram i64")
    return out


def sort_radix_signed(a: np.ndarray) -> np.ndarray:
    """SortRadixLsd(this int[] / long[]) in place."""
    f = lib().hpcs_oracle_sort_radix_lsd_i32 if a.dtype == np.int32 else lib().hpcs_oracle_sort_radix_lsd_i64
    _check(f(_ptr(a), a.size), "sort signed")
    return a
