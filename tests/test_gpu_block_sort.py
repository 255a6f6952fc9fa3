"""The one-workgroup sort (HPCS_OPT_PASS_ALGO = 3, and the default below
16384 4-byte / 8192 8-byte keys): the whole array sorted in LDS in one
launch.  Output against the oracle's restatements, bit for bit, at the
sizes around every limit, for every key kind and for key-value pairs."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
SIZES4 = (2, 3, 63, 64, 65, 1000, 4095, 4096, 4097, 8193, 16383, 16384, 16385)
SIZES8 = (2, 3, 65, 2047, 2048, 2049, 8191, 8192, 8193)


@pytest.fixture(params=["explicit", "default"])
def ctxb(request):
    ctx = N.new_context(0)
    if request.param == "explicit":
        N.set_option(N.HPCS_OPT_PASS_ALGO, 3, ctx)
    yield ctx
    N.destroy_context(ctx)


def _sort(ctx, a: np.ndarray) -> None:
    fn = {np.dtype(np.uint32): "hpcs_sort_radix_u32", np.dtype(np.uint64): "hpcs_sort_radix_u64",
          np.dtype(np.float32): "hpcs_sort_radix_f32", np.dtype(np.float64): "hpcs_sort_radix_f64",
          np.dtype(np.int32): "hpcs_sort_radix_i32", np.dtype(np.int64): "hpcs_sort_radix_i64"}[a.dtype]
    N.check(getattr(N.lib(), fn)(ctx, a.ctypes.data_as(P), a.size))


def test_plan_of_small_arrays(ctxb):
    info = N.PlanInfo()
    N.check(N.lib().hpcs_plan_query(ctxb, N.HPCS_DTYPE_U32, 4096, 0, ctypes.byref(info)))
    assert (info.tiles, info.block_threads, info.tile_keys, info.scratch_bytes) == (1, 256, 4096, 0)
    N.check(N.lib().hpcs_plan_query(ctxb, N.HPCS_DTYPE_U32, 16384, 0, ctypes.byref(info)))
    # the one-workgroup form up to 16384 4-byte keys, asked for or by size
    assert (info.tiles, info.block_threads, info.tile_keys, info.scratch_bytes) == (1, 1024, 16384, 0)
    N.check(N.lib().hpcs_plan_query(ctxb, N.HPCS_DTYPE_U32, 16385, 0, ctypes.byref(info)))
    assert info.tiles > 1
    N.check(N.lib().hpcs_plan_query(ctxb, N.HPCS_DTYPE_U64, 2048, 1, ctypes.byref(info)))
    assert (info.tiles, info.block_threads, info.tile_keys) == (1, 256, 2048)
    N.check(N.lib().hpcs_plan_query(ctxb, N.HPCS_DTYPE_U64, 8193, 0, ctypes.byref(info)))
    assert info.tiles > 1


@pytest.mark.parametrize("n", SIZES4)
def test_four_byte_keys(oracle, ctxb, n):
    for dist in ("uniform", "low8", "sorted", "constant"):
        a = datagen.keys_u32(n, 2000 + n, dist)
        ra = a.copy()
        oracle.sort_radix_lsd(ra)
        _sort(ctxb, a)
        assert np.array_equal(a, ra), dist
    x = datagen.floats_finite(n, 2100 + n)
    rx = x.copy()
    oracle.sort_radix_float(rx)
    _sort(ctxb, x)
    assert np.array_equal(x.view(np.uint32), rx.view(np.uint32))
    s = datagen.keys_u32(n, 2200 + n).view(np.int32).copy()
    rs = s.copy()
    oracle.sort_radix_signed(rs)
    _sort(ctxb, s)
    assert np.array_equal(s, rs)
    k = datagen.keys_u32(n, 2300 + n) & np.uint32(0xF0F)         # many ties: stability
    v = np.arange(n, dtype=np.uint32)
    rk, rv = k.copy(), v.copy()
    oracle.sort_radix_pairs(rk, rv)
    N.check(N.lib().hpcs_sort_radix_pairs_u32_u32(ctxb, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
    assert np.array_equal(k, rk) and np.array_equal(v, rv)


@pytest.mark.parametrize("n", SIZES8)
def test_eight_byte_keys(oracle, ctxb, n):
    a = datagen.keys_u64(n, 2400 + n)
    ra = a.copy()
    oracle.sort_radix_lsd(ra)
    _sort(ctxb, a)
    assert np.array_equal(a, ra)
    x = datagen.doubles_special(n, 2500 + n)
    rx = x.copy()
    oracle.sort_radix_float(rx)
    _sort(ctxb, x)
    assert np.array_equal(x.view(np.uint64), rx.view(np.uint64))
    s = datagen.keys_u64(n, 2600 + n).view(np.int64).copy()
    rs = s.copy()
    oracle.sort_radix_signed(rs)
    _sort(ctxb, s)
    assert np.array_equal(s, rs)
    k = datagen.keys_u64(n, 2700 + n) & np.uint64(0xFF000000000000FF)
    v = np.arange(n, dtype=np.uint32)
    order = np.argsort(k, kind="stable").astype(np.uint32)
    N.check(N.lib().hpcs_sort_radix_pairs_u64_u32(ctxb, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
    assert np.array_equal(v, order)
