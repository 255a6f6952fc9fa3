"""GPU parity of the batched segmented sort's tuning knobs: every ranking of
the large segments' scatter (HPCS_OPT_SEG_RANK 0-3), every tile shape of
their passes (HPCS_OPT_SEG_SHAPE: 256 x 8-32 and 512 x 16-32 keys) and both
limits of the segments sorted in LDS (HPCS_OPT_SEG_SMALL 4096 / 16384, the
latter with 1024-thread workgroups) against the oracle's per-segment sort
(SortRadixLsd(uint[], startIndex, length) of the Experimental class,
RadixSortLsd.cs:1409-1457, once per segment).  Bit-exact (integer work).

Segment lengths cover every path: empty and single keys, the small list
(<= 4096), the medium list (4097-16384), large segments whose last tile is
partial and exact multiples of the 4096-key tile; key distributions include
low-8-bit keys, a sorted run and a constant run, so passes where a tile's
keys share a digit run too."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    N.context(0)
    ctx = N.new_context(0)
    yield ctx
    N.destroy_context(ctx)


def _batch(seed):
    rng = np.random.default_rng(seed)
    lens = list(datagen.segment_lengths(180, seed, 1, 1 << 17))
    lens += [0, 1, 2, 4095, 4096, 4097, 8192, 12000, 16383, 16384, 16385, 20480, 4096 * 7 + 1, 65536]
    lens = np.array(lens, dtype=np.int64)
    rng.shuffle(lens)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64) + 5   # five untouched keys before
    n = int(begin[-1]) + 7                                                # and seven after
    a = datagen.keys_u32(n, seed + 1)
    a[: n // 4] &= np.uint32(0xFF)                                        # low-8 keys
    a[n // 4: n // 3] = np.arange(n // 3 - n // 4, dtype=np.uint32)       # a sorted run
    a[n // 2: n // 2 + 30000] = np.uint32(0xC0FFEE)                       # a constant run
    return a, begin


@pytest.mark.parametrize("small", [4096, 16384])
@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_segments_every_rank_and_lds_limit(oracle, sctx, rank, small):
    import torch
    a, begin = _batch(0x5E600 + rank)
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    N.set_option(N.HPCS_OPT_SEG_RANK, rank, sctx)
    N.set_option(N.HPCS_OPT_SEG_SMALL, small, sctx)
    t = torch.from_numpy(a.view(np.int32).copy()).cuda()
    b = torch.from_numpy(begin).cuda()
    st = torch.cuda.current_stream().cuda_stream
    print(f"segments rank {rank} small {small}: {begin.size - 1} segments, {a.size} keys", flush=True)
    N.check(N.lib().hpcs_dev_sort_segments_u32(sctx, ctypes.c_void_p(t.data_ptr()), a.size,
                                               ctypes.c_void_p(b.data_ptr()), begin.size - 1, ctypes.c_void_p(st)))
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ref)


SEG_SHAPES = [0, 1, 2, 3, 4, 7, 8, 9]


@pytest.mark.parametrize("shape", SEG_SHAPES)
@pytest.mark.parametrize("rank", [0, 2])
def test_segments_every_tile_shape(oracle, sctx, shape, rank):
    import torch
    a, begin = _batch(0x5E700 + shape)
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    N.set_option(N.HPCS_OPT_SEG_SHAPE, shape, sctx)
    N.set_option(N.HPCS_OPT_SEG_RANK, rank, sctx)
    N.set_option(N.HPCS_OPT_SEG_SMALL, 4096, sctx)
    try:
        t = torch.from_numpy(a.view(np.int32).copy()).cuda()
        b = torch.from_numpy(begin).cuda()
        st = torch.cuda.current_stream().cuda_stream
        print(f"segments shape {shape} rank {rank}: {begin.size - 1} segments, {a.size} keys", flush=True)
        N.check(N.lib().hpcs_dev_sort_segments_u32(sctx, ctypes.c_void_p(t.data_ptr()), a.size,
                                                   ctypes.c_void_p(b.data_ptr()), begin.size - 1,
                                                   ctypes.c_void_p(st)))
        torch.cuda.synchronize()
        assert np.array_equal(t.cpu().numpy().view(np.uint32), ref)
    finally:
        N.set_option(N.HPCS_OPT_SEG_SHAPE, 0, sctx)


def test_segment_options_reject_bad_values(sctx):
    with pytest.raises(N.HpcsError):
        N.set_option(N.HPCS_OPT_SEG_SMALL, 8192, sctx)
    with pytest.raises(N.HpcsError):
        N.set_option(N.HPCS_OPT_SEG_RANK, 4, sctx)
    for bad_shape in (5, 6, 10, -1):          # 1024-thread rows, 8-byte-only row, out of the table
        with pytest.raises(N.HpcsError):
            N.set_option(N.HPCS_OPT_SEG_SHAPE, bad_shape, sctx)


def test_segments_table_rows_drive_the_knobs(oracle, tmp_path):
    """A table whose `segs` rows pick LDS-OR ranking, the 4096-key limit and
    256x8 tiles for small batches and ballots, 16384 and 256x20 tiles for
    large ones sorts both batches as the oracle does."""
    import torch
    ctx = N.new_context(0)
    try:
        tbl = tmp_path / "segs.txt"
        tbl.write_text("segs 400000 4096 1 4\nsegs 100000000 16384 0 8\n")
        N.check(N.lib().hpcs_ctx_load_tuning(ctx, str(tbl).encode()))
        for seed, cap in ((11, 1 << 12), (12, 1 << 17)):
            lens = datagen.segment_lengths(40, seed, 1, cap)
            begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
            a = datagen.keys_u32(int(begin[-1]), seed)
            ref = a.copy()
            oracle.sort_segments_u32(ref, begin)
            t = torch.from_numpy(a.view(np.int32).copy()).cuda()
            b = torch.from_numpy(begin).cuda()
            N.check(N.lib().hpcs_dev_sort_segments_u32(ctx, ctypes.c_void_p(t.data_ptr()), a.size,
                                                       ctypes.c_void_p(b.data_ptr()), lens.size,
                                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
            torch.cuda.synchronize()
            assert np.array_equal(t.cpu().numpy().view(np.uint32), ref), seed
    finally:
        N.destroy_context(ctx)
