"""The block-sort + merge form (HPCS_OPT_PASS_ALGO = 5, up to 2^20 keys):
runs of 4096 4-byte / 2048 8-byte keys sorted in LDS by one workgroup each,
then stable merge passes.  Output against the oracle's restatements, bit for
bit, at sizes around the run length, the merge tile and powers of two, for
every key kind, for key-value pairs (stability: narrow keys with the index
as value) and for device-resident calls; a chunk whose keys are all equal
(no radix pass runs in it) must still reach the merge buffers."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
SIZES4 = (2, 4097, 8191, 8192, 12289, 65536 + 77, 1 << 18, (1 << 20) - 3, 1 << 20)
SIZES8 = (2, 2049, 4095, 4096, 6145, 40000, 1 << 17, (1 << 20) - 1)


@pytest.fixture(scope="module")
def ctxm():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    N.context(0)
    ctx = N.new_context(0)
    N.set_option(N.HPCS_OPT_PASS_ALGO, 5, ctx)
    yield ctx
    N.destroy_context(ctx)


def _sort(ctx, a: np.ndarray) -> None:
    fn = {np.dtype(np.uint32): "hpcs_sort_radix_u32", np.dtype(np.uint64): "hpcs_sort_radix_u64",
          np.dtype(np.float32): "hpcs_sort_radix_f32", np.dtype(np.float64): "hpcs_sort_radix_f64",
          np.dtype(np.int32): "hpcs_sort_radix_i32", np.dtype(np.int64): "hpcs_sort_radix_i64"}[a.dtype]
    N.check(getattr(N.lib(), fn)(ctx, a.ctypes.data_as(P), a.size))


def test_plan_of_the_merge_form(ctxm):
    info = N.PlanInfo()
    N.check(N.lib().hpcs_plan_query(ctxm, N.HPCS_DTYPE_U32, 1 << 16, 0, ctypes.byref(info)))
    assert (info.tiles, info.block_threads, info.tile_keys) == (16, 256, 4096)
    assert info.ideal_bytes == (1 << 16) * 5 * 2 * 4           # the runs, then four merge passes
    N.check(N.lib().hpcs_plan_query(ctxm, N.HPCS_DTYPE_U64, 1 << 14, 1, ctypes.byref(info)))
    assert (info.tiles, info.tile_keys, info.scratch_bytes >= (1 << 14) * 12) == (8, 2048, True)
    n = (1 << 20) + 1                              # beyond 2^20 keys: the LSD passes (choice by size)
    N.check(N.lib().hpcs_plan_query(ctxm, N.HPCS_DTYPE_U32, n, 0, ctypes.byref(info)))
    assert (info.passes, info.ideal_bytes) == (4, n * 36)


@pytest.mark.parametrize("n", SIZES4)
def test_four_byte_keys(oracle, ctxm, n):
    print(f"merge form, 4-byte keys, n={n}", flush=True)
    for dist in ("uniform", "low8", "sorted", "constant"):
        a = datagen.keys_u32(n, 3000 + n, dist)
        ra = a.copy()
        oracle.sort_radix_lsd(ra)
        _sort(ctxm, a)
        assert np.array_equal(a, ra), dist
    x = datagen.floats_finite(n, 3100 + n)
    rx = x.copy()
    oracle.sort_radix_float(rx)
    _sort(ctxm, x)
    assert np.array_equal(x.view(np.uint32), rx.view(np.uint32))
    s = datagen.keys_u32(n, 3200 + n).view(np.int32).copy()
    rs = s.copy()
    oracle.sort_radix_signed(rs)
    _sort(ctxm, s)
    assert np.array_equal(s, rs)
    k = datagen.keys_u32(n, 3300 + n) & np.uint32(0xF0F)         # many ties: stability
    v = np.arange(n, dtype=np.uint32)
    rk, rv = k.copy(), v.copy()
    oracle.sort_radix_pairs(rk, rv)
    N.check(N.lib().hpcs_sort_radix_pairs_u32_u32(ctxm, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
    assert np.array_equal(k, rk) and np.array_equal(v, rv)


@pytest.mark.parametrize("n", SIZES8)
def test_eight_byte_keys(oracle, ctxm, n):
    print(f"merge form, 8-byte keys, n={n}", flush=True)
    a = datagen.keys_u64(n, 3400 + n)
    ra = a.copy()
    oracle.sort_radix_lsd(ra)
    _sort(ctxm, a)
    assert np.array_equal(a, ra)
    x = datagen.doubles_special(n, 3500 + n)
    rx = x.copy()
    oracle.sort_radix_float(rx)
    _sort(ctxm, x)
    assert np.array_equal(x.view(np.uint64), rx.view(np.uint64))
    s = datagen.keys_u64(n, 3600 + n).view(np.int64).copy()
    rs = s.copy()
    oracle.sort_radix_signed(rs)
    _sort(ctxm, s)
    assert np.array_equal(s, rs)
    k = datagen.keys_u64(n, 3700 + n) & np.uint64(0xFF000000000000FF)
    v = np.arange(n, dtype=np.uint32)
    order = np.argsort(k, kind="stable").astype(np.uint32)
    N.check(N.lib().hpcs_sort_radix_pairs_u64_u32(ctxm, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
    assert np.array_equal(v, order)


def test_equal_chunks_and_device_calls(oracle, ctxm):
    """Runs whose keys are all equal (no radix pass runs in them) next to
    ordinary runs, through the device entry points with values."""
    import torch
    n = 5 * 4096 + 123
    a = datagen.keys_u32(n, 3800)
    a[4096:2 * 4096] = 7                          # run 1: one value
    a[3 * 4096:4 * 4096] = 0xFFFFFFFF             # run 3: one value
    v = np.arange(n, dtype=np.uint32)
    rk, rv = a.copy(), v.copy()
    oracle.sort_radix_pairs(rk, rv)
    tk = torch.from_numpy(a.view(np.int32).copy()).cuda()
    tv = torch.from_numpy(v.view(np.int32).copy()).cuda()
    st = torch.cuda.current_stream().cuda_stream
    N.check(N.lib().hpcs_dev_sort_radix_pairs_u32_u32(ctxm, P(tk.data_ptr()), P(tv.data_ptr()), n, P(st)))
    torch.cuda.synchronize()
    assert np.array_equal(tk.cpu().numpy().view(np.uint32), rk)
    assert np.array_equal(tv.cpu().numpy().view(np.uint32), rv)
    b = datagen.keys_u64(3 * 2048 + 5, 3900)
    b[:2048] = 42
    rb = b.copy()
    oracle.sort_radix_lsd(rb)
    tb = torch.from_numpy(b.view(np.int64).copy()).cuda()
    N.check(N.lib().hpcs_dev_sort_radix_u64(ctxm, P(tb.data_ptr()), b.size, P(st)))
    torch.cuda.synchronize()
    assert np.array_equal(tb.cpu().numpy().view(np.uint64), rb)
