"""Sorts with 11-bit digits (HPCS_OPT_DIGIT_BITS = 11): three passes for
4-byte keys, six for 8-byte keys, in the single-pass look-back form.  The
oracle is the restatement of SortRadixLsdNbit (RadixSortLsd.cs:1616-1661)
with 11-bit digits for unsigned keys, and the float / signed restatements
for the other key kinds; the output must match bit for bit."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
SIZES = (2, 3, 4095, 4096, 4097, 3 * 4096 + 5, 100003, (1 << 20) + 3)


@pytest.fixture
def ctx11():
    ctx = N.new_context(0)
    N.set_option(N.HPCS_OPT_DIGIT_BITS, 11, ctx)
    N.set_option(N.HPCS_OPT_PASS_ALGO, 0, ctx)
    yield ctx
    N.destroy_context(ctx)


def _sort(ctx, a: np.ndarray) -> None:
    fn = {np.dtype(np.uint32): "hpcs_sort_radix_u32", np.dtype(np.uint64): "hpcs_sort_radix_u64",
          np.dtype(np.float32): "hpcs_sort_radix_f32", np.dtype(np.float64): "hpcs_sort_radix_f64",
          np.dtype(np.int32): "hpcs_sort_radix_i32", np.dtype(np.int64): "hpcs_sort_radix_i64"}[a.dtype]
    N.check(getattr(N.lib(), fn)(ctx, a.ctypes.data_as(P), a.size))


def test_plan_reports_eleven_bit_digits(ctx11):
    info = N.PlanInfo()
    N.check(N.lib().hpcs_plan_query(ctx11, N.HPCS_DTYPE_U32, 1 << 16, 0, ctypes.byref(info)))
    assert (info.digit_bits, info.passes, info.tile_keys) == (11, 3, 4096)
    assert info.ideal_bytes == (1 << 16) * (4 + 3 * 8)
    N.check(N.lib().hpcs_plan_query(ctx11, N.HPCS_DTYPE_F64, 1 << 16, 0, ctypes.byref(info)))
    assert (info.digit_bits, info.passes) == (11, 6)
    # values, or the reduce-then-scan form, keep 8-bit digits
    N.check(N.lib().hpcs_plan_query(ctx11, N.HPCS_DTYPE_U64, 1 << 16, 1, ctypes.byref(info)))
    assert (info.digit_bits, info.passes) == (8, 8)
    with pytest.raises(N.HpcsError):
        N.set_option(N.HPCS_OPT_DIGIT_BITS, 10, ctx11)


@pytest.mark.parametrize("n", SIZES)
def test_unsigned_keys_vs_nbit_oracle(oracle, ctx11, n):
    a = datagen.keys_u32(n, 1100 + n)
    ra = a.copy()
    oracle.sort_radix_lsd_nbit(ra, 11)
    _sort(ctx11, a)
    assert np.array_equal(a, ra)
    b = datagen.keys_u64(n, 1200 + n)
    rb = b.copy()
    oracle.sort_radix_lsd_nbit(rb, 11)
    _sort(ctx11, b)
    assert np.array_equal(b, rb)


@pytest.mark.parametrize("n", (4097, 100003))
def test_float_and_signed_keys(oracle, ctx11, n):
    x = datagen.floats_finite(n, 1300 + n)
    rx = x.copy()
    oracle.sort_radix_float(rx)
    _sort(ctx11, x)
    assert np.array_equal(x.view(np.uint32), rx.view(np.uint32))
    y = datagen.doubles_special(n, 1400 + n)
    ry = y.copy()
    oracle.sort_radix_float(ry)
    _sort(ctx11, y)
    assert np.array_equal(y.view(np.uint64), ry.view(np.uint64))
    for dt, gen in ((np.int32, datagen.keys_u32), (np.int64, datagen.keys_u64)):
        s = gen(n, 1500 + n).view(dt).copy()
        rs = s.copy()
        oracle.sort_radix_signed(rs)
        _sort(ctx11, s)
        assert np.array_equal(s, rs), dt


@pytest.mark.parametrize("mask", (0x00000000, 0x000007FF, 0x003FF800, 0xFFC00000, 0xFFC007FF))
def test_passes_that_move_nothing(oracle, ctx11, mask):
    """Digits with one value everywhere are skipped; the result lands in the
    caller's buffer whatever the number of moving passes."""
    n = 50001
    a = datagen.keys_u32(n, 1600) & np.uint32(mask)
    ra = a.copy()
    oracle.sort_radix_lsd(ra)
    _sort(ctx11, a)
    assert np.array_equal(a, ra)
    x = (datagen.keys_u32(n, 1601) & np.uint32(mask)).view(np.float32)
    x = np.where(np.isnan(x), np.float32(0), x).astype(np.float32)
    rx = x.copy()
    oracle.sort_radix_float(rx)
    _sort(ctx11, x)
    assert np.array_equal(x.view(np.uint32), rx.view(np.uint32))


def test_look_back_fallback(oracle, ctx11):
    """With no look-back rounds allowed, every wait counts the earlier tile
    itself: same output."""
    N.set_option(N.HPCS_OPT_FALLBACK_ROUNDS, 0, ctx11)
    for n in (3 * 4096 + 5, 40 * 4096):
        a = datagen.keys_u32(n, 1700 + n)
        ra = a.copy()
        oracle.sort_radix_lsd(ra)
        _sort(ctx11, a)
        assert np.array_equal(a, ra)
        b = datagen.keys_u64(n, 1800 + n)
        rb = b.copy()
        oracle.sort_radix_lsd(rb)
        _sort(ctx11, b)
        assert np.array_equal(b, rb)


def test_tuning_row_selects_eleven_bits(oracle, tmp_path):
    ctx = N.new_context(0)
    try:
        p = tmp_path / "tune.txt"
        p.write_text("u32 1048576 0 0 0 11\nu32 268435456 1 0 1\n")
        N.load_tuning(str(p), ctx)
        info = N.PlanInfo()
        N.check(N.lib().hpcs_plan_query(ctx, N.HPCS_DTYPE_U32, 1 << 20, 0, ctypes.byref(info)))
        assert (info.digit_bits, info.passes) == (11, 3)
        N.check(N.lib().hpcs_plan_query(ctx, N.HPCS_DTYPE_U32, 1 << 21, 0, ctypes.byref(info)))
        assert (info.digit_bits, info.passes) == (8, 4)
        N.set_option(N.HPCS_OPT_DIGIT_BITS, 8, ctx)            # an explicit option beats the table
        N.check(N.lib().hpcs_plan_query(ctx, N.HPCS_DTYPE_U32, 1 << 20, 0, ctypes.byref(info)))
        assert info.digit_bits == 8
        N.load_tuning(str(p), ctx)
        a = datagen.keys_u32(777777, 1900)
        ra = a.copy()
        oracle.sort_radix_lsd(ra)
        N.check(N.lib().hpcs_sort_radix_u32(ctx, a.ctypes.data_as(P), a.size))
        assert np.array_equal(a, ra)
    finally:
        N.destroy_context(ctx)


def test_python_mirror_11bit_and_nbit_sorts(oracle):
    """Algorithm.SortRadixLsd11bit / SortRadixLsdNbit (RadixSortLsd.cs:1484-1546,
    1616-1661, 1839-1882) against the oracle's SortRadixLsdNbit restatement:
    whole arrays, sub-ranges (only the range changes), several widths."""
    from hpcsharp_amd.algorithm import Algorithm
    a = datagen.keys_u32(200003, 3100)
    ra = a.copy()
    oracle.sort_radix_lsd_nbit(ra, 11)
    assert Algorithm.SortRadixLsd11bit(a) is a
    assert np.array_equal(a, ra)
    b = datagen.keys_u32(100000, 3101)
    rb = b.copy()
    rb[1000:61000] = np.sort(rb[1000:61000])
    Algorithm.SortRadixLsd11bit(b, 1000, 60000)
    assert np.array_equal(b, rb)
    for bits in (1, 5, 8, 11, 16, 32):
        c = datagen.keys_u32(50001, 3200 + bits)
        rc = c.copy()
        rc[7:50001 - 9] = np.sort(rc[7:50001 - 9])
        Algorithm.SortRadixLsdNbit(c, 7, 50001 - 16, bits)
        assert np.array_equal(c, rc), bits
    for bits in (11, 13, 64):
        d = datagen.keys_u64(40009, 3300 + bits)
        rd = d.copy()
        if bits <= 16:
            oracle.sort_radix_lsd_nbit(rd, bits)
        else:
            rd.sort()
        Algorithm.SortRadixLsdNbit(d, 0, d.size, bits)
        assert np.array_equal(d, rd), bits
