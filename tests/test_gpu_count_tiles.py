"""Tile counts of several adjacent tiles per workgroup (HPCS_OPT_COUNT_TILES):
the reduce-then-scan sort gives the oracle's output (RadixSortLsd.cs:127-202
and :359-384 restated) on the lean and the full pass sequences, with ragged
last tiles, tile counts not divisible by the tiles per workgroup, and passes
that move nothing."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p


def test_count_tiles_option_range():
    c = N.new_context(0)
    try:
        for bad in (0, 3, 8):
            with pytest.raises(N.HpcsError):
                N.set_option(N.HPCS_OPT_COUNT_TILES, bad, c)
    finally:
        N.destroy_context(c)


@pytest.mark.parametrize("tpw", [2, 4])
@pytest.mark.parametrize("n", [(1 << 20) + 4097, 3 * (1 << 20) + 3 * 4096 + 5])
def test_count_tiles_sorts(oracle, tpw, n):
    c = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, 1, c)
        N.set_option(N.HPCS_OPT_COUNT_TILES, tpw, c)
        base = datagen.keys_u32(n, 5150 + tpw)
        for mask in (0xFFFFFFFF, 0x00FF00FF, 0xFFFF0000, 0x0):
            a = base & np.uint32(mask)
            ra = a.copy()
            oracle.sort_radix_lsd(ra)
            print("count tiles u32", n, tpw, hex(mask), flush=True)
            N.check(N.lib().hpcs_sort_radix_u32(c, a.ctypes.data_as(P), n))
            assert np.array_equal(a, ra), (n, tpw, hex(mask))
        f = datagen.keys_u32(n, 5160 + tpw)
        f = np.where(((f >> np.uint32(23)) & np.uint32(0xFF)) == 0xFF, f & np.uint32(0xFF7FFFFF), f)
        f = f.view(np.float32).copy()
        rf = f.copy()
        oracle.sort_radix_float(rf)
        print("count tiles f32", n, tpw, flush=True)
        N.check(N.lib().hpcs_sort_radix_f32(c, f.ctypes.data_as(P), n))
        assert np.array_equal(f.view(np.uint32), rf.view(np.uint32))
        b = datagen.keys_u64(n, 5170 + tpw) & np.uint64(0xFFFF00FF00FFFFFF)
        rb = b.copy()
        oracle.sort_radix_lsd(rb)
        print("count tiles u64", n, tpw, flush=True)
        N.check(N.lib().hpcs_sort_radix_u64(c, b.ctypes.data_as(P), n))
        assert np.array_equal(b, rb)
        k = datagen.keys_u32(n, 5180 + tpw) & np.uint32(0xFFF)
        v = np.arange(n, dtype=np.uint32)
        order = np.argsort(k, kind="stable").astype(np.uint32)
        print("count tiles pairs32", n, tpw, flush=True)
        N.check(N.lib().hpcs_sort_radix_pairs_u32_u32(c, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
        assert np.array_equal(v, order)
    finally:
        N.destroy_context(c)
