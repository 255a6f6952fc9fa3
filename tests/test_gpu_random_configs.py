"""Parity over many option combinations: fixed-seed random cases, each a key
type, a size (log-uniform, 2 .. 2^22), a distribution and explicitly set
options (pass algorithm, including the MSD split from case 80, tile shape,
ranking, digit width).  The expected
result is numpy's sort (keys, in the order the reference's key transforms
define) or its stable argsort (key-value pairs)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu

P = ctypes.c_void_p
TYPES = ["u32", "u64", "f32", "f64", "i32", "i64", "pairs32", "pairs64"]
DISTS = ["uniform", "lowbits", "few", "sorted", "reverse", "constant"]


def _keys(rng, kind, n):
    dist = DISTS[rng.integers(len(DISTS))]
    wide = kind in ("u64", "f64", "i64", "pairs64")
    x = datagen.keys_u64(n, int(rng.integers(1 << 30))) if wide else datagen.keys_u32(n, int(rng.integers(1 << 30)))
    if dist == "lowbits":
        x = x & x.dtype.type(0xFFF)
    elif dist == "few":
        x = (x % x.dtype.type(7)) * x.dtype.type(0x01010101)
    elif dist == "sorted":
        x = np.sort(x)
    elif dist == "reverse":
        x = np.sort(x)[::-1].copy()
    elif dist == "constant":
        x = np.full(n, x[0] if n else 0, dtype=x.dtype)
    if kind == "f32":
        x = np.where(((x >> np.uint32(23)) & np.uint32(0xFF)) == 0xFF, x & np.uint32(0xFF7FFFFF), x).view(np.float32)
    elif kind == "f64":
        e = (x >> np.uint64(52)) & np.uint64(0x7FF)
        x = np.where(e == 0x7FF, x & np.uint64(0xFFEFFFFFFFFFFFFF), x).view(np.float64)
    elif kind == "i32":
        x = x.view(np.int32)
    elif kind == "i64":
        x = x.view(np.int64)
    return np.ascontiguousarray(x), dist


def _float_order(x):
    """Ascending order of the reference's float transform: numeric, -0.0
    before +0.0 (NaN-free input)."""
    bits = x.view(np.uint32 if x.dtype == np.float32 else np.uint64)
    sign = bits >> np.uint64(8 * x.itemsize - 1) if x.itemsize == 8 else bits >> np.uint32(31)
    return np.lexsort((sign == 0, x))


@pytest.mark.parametrize("case", range(120))
def test_random_option_combinations(case):
    rng = np.random.default_rng(0xC0FFEE + case)
    kind = TYPES[case % len(TYPES)]
    n = int(np.exp(rng.uniform(np.log(2), np.log(1 << 22))))
    # cases from 80 also draw the MSD split (pass algorithm 4)
    algo = [2, 0, 1, 3][int(rng.integers(4))] if case < 80 else [2, 0, 1, 3, 4][int(rng.integers(5))]
    rank = int(rng.integers(4))      # ballot, LDS-OR, and the two mixes
    four = kind in ("u32", "f32", "i32")
    # the rows of the tile-shape table that serve the key type
    shape = int(rng.choice([0, 1, 2, 3, 4, 5, 7, 8, 9] if four else [0, 4, 6] if kind != "pairs64" else [0, 4]))
    bits = 11 if (not kind.startswith("pairs") and rng.random() < 0.25) else 8
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, algo, ctx)
        N.set_option(N.HPCS_OPT_RANK, rank, ctx)
        N.set_option(N.HPCS_OPT_TILE_SHAPE, shape, ctx)
        N.set_option(N.HPCS_OPT_DIGIT_BITS, bits, ctx)
        x, dist = _keys(rng, kind, n)
        tag = (kind, n, dist, algo, rank, shape, bits)
        lib = N.lib()
        if kind.startswith("pairs"):
            k = x
            v = np.arange(n, dtype=np.uint32)
            order = np.argsort(k, kind="stable").astype(np.uint32)
            fn = lib.hpcs_sort_radix_pairs_u32_u32 if kind == "pairs32" else lib.hpcs_sort_radix_pairs_u64_u32
            N.check(fn(ctx, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
            assert np.array_equal(v, order), tag
            return
        want = x[_float_order(x)] if kind in ("f32", "f64") else np.sort(x, kind="stable")
        fn = getattr(lib, "hpcs_sort_radix_" + kind)
        N.check(fn(ctx, x.ctypes.data_as(P), n))
        assert np.array_equal(x.view(np.uint8), want.view(np.uint8)), tag
    finally:
        N.destroy_context(ctx)
