"""The line-carrying scatter (HPCS_OPT_LINE_CHUNK): a workgroup scatters a
chunk of consecutive tiles and writes each digit's output in whole lines.
Its output must equal the oracle's (RadixSortLsd.cs:127-202 restated) for
4-byte keys of every kind, on the lean (few tiles) and the full
reduce-then-scan pass sequences, with ragged last tiles, passes that move
nothing, chunks of one tile, chunks longer than the array, and chunk counts
that do not divide the tile count.  Sorts the option does not cover (8-byte
keys, key-value pairs) ignore it."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
# 2^20 + 4097: 257 tiles of 4096 keys (lean path, ragged last tile);
# 3 * 2^20 + 4097: 769 tiles (full reduce-then-scan path)
SIZES = (1 << 20, (1 << 20) + 4097, 3 * (1 << 20) + 4097)


@pytest.fixture
def ctx():
    c = N.new_context(0)
    N.set_option(N.HPCS_OPT_PASS_ALGO, 1, c)
    yield c
    N.destroy_context(c)


def test_line_chunk_option_range(ctx):
    for bad in (-1, 257):
        with pytest.raises(N.HpcsError):
            N.set_option(N.HPCS_OPT_LINE_CHUNK, bad, ctx)
    N.set_option(N.HPCS_OPT_LINE_CHUNK, 256, ctx)
    N.set_option(N.HPCS_OPT_LINE_CHUNK, 0, ctx)


@pytest.mark.parametrize("chunk", [1, 3, 16, 256])
@pytest.mark.parametrize("n", SIZES)
def test_lines_u32_masks(oracle, ctx, chunk, n):
    N.set_option(N.HPCS_OPT_LINE_CHUNK, chunk, ctx)
    base = datagen.keys_u32(n, 4242 + chunk)
    for mask in (0xFFFFFFFF, 0xFFFFFF00, 0x00FF00FF, 0xFF000000, 0x00000007, 0x0):
        a = base & np.uint32(mask)
        ra = a.copy()
        oracle.sort_radix_lsd(ra)
        print("lines u32", n, chunk, hex(mask), flush=True)
        N.check(N.lib().hpcs_sort_radix_u32(ctx, a.ctypes.data_as(P), n))
        assert np.array_equal(a, ra), (n, chunk, hex(mask))


@pytest.mark.parametrize("chunk", [2, 7, 32])
def test_lines_float_and_signed(oracle, ctx, chunk):
    N.set_option(N.HPCS_OPT_LINE_CHUNK, chunk, ctx)
    n = SIZES[2]
    x = datagen.keys_u32(n, 77 + chunk)
    # finite float bit patterns, with +-0.0 and subnormals mixed in
    x = np.where(((x >> np.uint32(23)) & np.uint32(0xFF)) == 0xFF, x & np.uint32(0xFF7FFFFF), x)
    x[::97] = np.uint32(0x80000000)
    x[1::101] = np.uint32(0)
    x[2::103] &= np.uint32(0x807FFFFF)
    f = x.view(np.float32).copy()
    rf = f.copy()
    oracle.sort_radix_float(rf)
    print("lines f32", n, chunk, flush=True)
    N.check(N.lib().hpcs_sort_radix_f32(ctx, f.ctypes.data_as(P), n))
    assert np.array_equal(f.view(np.uint32), rf.view(np.uint32))
    i = datagen.keys_u32(n, 78 + chunk).view(np.int32).copy()
    want = np.sort(i, kind="stable")
    print("lines i32", n, chunk, flush=True)
    N.check(N.lib().hpcs_sort_radix_i32(ctx, i.ctypes.data_as(P), n))
    assert np.array_equal(i, want)


def test_lines_ignored_by_other_sorts(oracle, ctx):
    N.set_option(N.HPCS_OPT_LINE_CHUNK, 8, ctx)
    n = SIZES[2]
    b = datagen.keys_u64(n, 91)
    rb = b.copy()
    oracle.sort_radix_lsd(rb)
    print("lines option, u64", n, flush=True)
    N.check(N.lib().hpcs_sort_radix_u64(ctx, b.ctypes.data_as(P), n))
    assert np.array_equal(b, rb)
    k = datagen.keys_u32(n, 92) & np.uint32(0xFFFF)
    v = np.arange(n, dtype=np.uint32)
    order = np.argsort(k, kind="stable").astype(np.uint32)
    print("lines option, pairs32", n, flush=True)
    N.check(N.lib().hpcs_sort_radix_pairs_u32_u32(ctx, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
    assert np.array_equal(v, order)
