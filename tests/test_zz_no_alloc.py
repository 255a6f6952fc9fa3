"""The no-allocation sort (hpcs_dev_sort_radix_u32_noalloc): its scratch is
the library's static device array, and its output equals
hpcs_dev_sort_radix_u32's bit for bit, on consecutive calls with a smaller
second input (the array keeps the previous call's contents), on every pass
form and on the largest size it takes; larger inputs are rejected."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import _native as N
from hpcsharp_amd import datagen

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p


def _pair(torch, lib, ctx, h):
    """(plain entry point's output, no-allocation entry point's output)."""
    st = torch.cuda.current_stream()
    src = torch.from_numpy(h.view(np.int32)).cuda()
    a, b = src.clone(), src.clone()
    N.check(lib.hpcs_dev_sort_radix_u32(ctx, P(a.data_ptr()), h.size, P(st.cuda_stream)))
    N.check(lib.hpcs_dev_sort_radix_u32_noalloc(ctx, P(b.data_ptr()), h.size, P(st.cuda_stream)))
    torch.cuda.synchronize()
    return a.cpu().numpy().view(np.uint32), b.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("algo", [2, 0, 1])
def test_noalloc_matches_the_plain_sort(algo):
    import torch
    lib = N.lib()
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, algo, ctx)
        # two calls in a row, the second smaller, then a one-workgroup size
        for n, seed, dist in ((3 * (1 << 20) + 5, 1, "uniform"), (100003, 2, "low8"), (5000, 3, "uniform")):
            h = datagen.keys_u32(n, 0x5EED0100 + seed, dist)
            plain, noalloc = _pair(torch, lib, ctx, h)
            assert np.array_equal(noalloc, plain), (algo, n)
            assert np.array_equal(noalloc, np.sort(h)), (algo, n)
    finally:
        N.destroy_context(ctx)


@pytest.mark.parametrize("shape", [0, 4, 5])
def test_noalloc_largest_size_every_tile(shape):
    """2^22 keys, the largest accepted, on the tile shapes that make the
    largest and the smallest scratch layouts."""
    import torch
    lib = N.lib()
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, 1, ctx)
        N.set_option(N.HPCS_OPT_TILE_SHAPE, shape, ctx)
        h = datagen.keys_u32(1 << 22, 0x5EED0200 + shape)
        plain, noalloc = _pair(torch, lib, ctx, h)
        assert np.array_equal(noalloc, plain)
        # a second, smaller call on the same static scratch
        h2 = datagen.keys_u32((1 << 21) + 77, 0x5EED0300 + shape, "sorted")[::-1].copy()
        plain2, noalloc2 = _pair(torch, lib, ctx, h2)
        assert np.array_equal(noalloc2, plain2) and np.array_equal(noalloc2, np.sort(h2))
    finally:
        N.destroy_context(ctx)


def test_noalloc_rejects_beyond_its_limit():
    import torch
    lib = N.lib()
    ctx = N.new_context(0)
    try:
        n = (1 << 22) + 1
        buf = torch.zeros(n, dtype=torch.int32, device="cuda")
        rc = lib.hpcs_dev_sort_radix_u32_noalloc(ctx, P(buf.data_ptr()), n, P(torch.cuda.current_stream().cuda_stream))
        assert rc == N.HPCS_E_UNSUPPORTED
        assert b"2^22" in lib.hpcs_last_error()
        assert lib.hpcs_dev_sort_radix_u32_noalloc(ctx, P(buf.data_ptr()), -1, None) == N.HPCS_E_INVALID_ARG
        assert lib.hpcs_dev_sort_radix_u32_noalloc(ctx, P(buf.data_ptr()), 1, None) == N.HPCS_OK
    finally:
        N.destroy_context(ctx)
