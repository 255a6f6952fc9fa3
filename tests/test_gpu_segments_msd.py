"""GPU parity: the batched segmented sort's MSD-first form
(HPCS_OPT_SEG_MSD): segments longer than the threshold take one segmented
pass by the top byte, then every (segment, top byte) bucket is sorted in LDS.
Bit-exact against the oracle's per-segment SortRadixLsd(uint[], startIndex,
length) of the Experimental class (RadixSortLsd.cs:1409-1457, once per
segment): keys alone carry no identity, so the unstable bucket sort gives the
same array.

Segments and keys reach every branch of the bucket stage:
- uniform keys: buckets of at most 5120 keys sorted in LDS;
- segments of 3 x 2^20 keys: buckets of ~12288 keys, the 1024-thread LDS
  sort of the overflow;
- a segment of 2^23 keys and segments whose keys share the top byte (keys
  below 2^24): buckets beyond 16384 keys, the three overflow passes;
- few distinct values: 12-bit bins beyond 256 keys (overflow);
- ascending, constant and low-zero keys, empty buckets and buckets of one
  key; segments at and around the threshold and beside the small, medium
  and plain large segments of the same call.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p


@pytest.fixture(scope="module")
def mctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = N.new_context(0)
    yield ctx
    torch.cuda.synchronize()
    N.destroy_context(ctx)


def _fill(a: np.ndarray, lo: int, hi: int, kind: int, seed: int) -> None:
    m = hi - lo
    if m <= 0:
        return
    r = datagen.keys_u32(m, seed)
    if kind == 1:
        r &= np.uint32(0xFFFFFF)                      # top byte constant: one bucket
    elif kind == 2:
        r = (r % 300).astype(np.uint32) * np.uint32(14000001)  # 300 values: crowded bins
    elif kind == 3:
        r = np.arange(m, dtype=np.uint32) * np.uint32(37)       # ascending, spread
    elif kind == 4:
        r = np.full(m, 0xA5A5A5A5, np.uint32)                  # constant
    elif kind == 5:
        r &= np.uint32(0xFFFFF000)                    # low 12 bits zero
    elif kind == 6:
        r = (r & np.uint32(0x0F0000FF)) | np.uint32(0x30000000)  # 16 top bytes, sparse bits
    a[lo:hi] = r


def _batch(seed: int, big: bool):
    rng = np.random.default_rng(seed)
    lens = list(datagen.segment_lengths(120, seed, 1, 1 << 18))
    lens += [0, 1, 4096, 4097, 16384, 16385, 16386, 20000, 65536, 65537, (1 << 20) + 3, 3 * (1 << 20) + 1]
    if big:
        lens += [1 << 23]
    lens = np.array(lens, dtype=np.int64)
    rng.shuffle(lens)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64) + 3   # three untouched keys before
    n = int(begin[-1]) + 5                                                # and five after
    a = datagen.keys_u32(n, seed + 1)
    for s in range(lens.size):
        _fill(a, int(begin[s]), int(begin[s + 1]), int(rng.integers(0, 7)) if s % 3 else 0, seed + 7 * s)
    return a, begin


def _run(ctx, a: np.ndarray, begin: np.ndarray) -> np.ndarray:
    import torch
    t = torch.from_numpy(a.view(np.int32).copy()).cuda()
    b = torch.from_numpy(begin).cuda()
    N.check(N.lib().hpcs_dev_sort_segments_u32(ctx, P(t.data_ptr()), a.size, P(b.data_ptr()), begin.size - 1,
                                               P(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("small", [4096, 16384])
@pytest.mark.parametrize("thr", [16385, 65536, 1 << 20])
def test_segments_msd_matches_oracle(oracle, mctx, thr, small):
    a, begin = _batch(0x3D0 + thr % 977 + small, big=False)
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    N.set_option(N.HPCS_OPT_SEG_SMALL, small, mctx)
    N.set_option(N.HPCS_OPT_SEG_MSD, thr, mctx)
    print(f"segments msd thr {thr} small {small}: {begin.size - 1} segments, {a.size} keys", flush=True)
    try:
        got = _run(mctx, a, begin)
    finally:
        N.set_option(N.HPCS_OPT_SEG_MSD, 0, mctx)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("rank,shape", [(0, 0), (1, 2), (3, 7)])
def test_segments_msd_large_overflow(oracle, mctx, rank, shape):
    """A segment of 2^23 keys (uniform: 32768 keys per bucket) takes the
    three overflow passes; every ranking and some tile shapes."""
    a, begin = _batch(0x8F1 + rank, big=True)
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    N.set_option(N.HPCS_OPT_SEG_RANK, rank, mctx)
    N.set_option(N.HPCS_OPT_SEG_SHAPE, shape, mctx)
    N.set_option(N.HPCS_OPT_SEG_MSD, 16385, mctx)
    print(f"segments msd overflow rank {rank} shape {shape}: {a.size} keys", flush=True)
    try:
        got = _run(mctx, a, begin)
    finally:
        N.set_option(N.HPCS_OPT_SEG_MSD, 0, mctx)
    assert np.array_equal(got, ref)


def test_segments_msd_bench_shape(mctx):
    """The benchmark's shape at a quarter of its count: 1024 segments of
    log-uniform lengths in [16, 2^20], uniform keys; np.sort per segment."""
    lens = datagen.segment_lengths(1024, 0x5EED, 16, 1 << 20)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a = datagen.keys_u32(int(begin[-1]), 0x5EEE)
    N.set_option(N.HPCS_OPT_SEG_MSD, 1 << 16, mctx)
    try:
        got = _run(mctx, a, begin)
    finally:
        N.set_option(N.HPCS_OPT_SEG_MSD, 0, mctx)
    for s in range(lens.size):
        lo, hi = int(begin[s]), int(begin[s + 1])
        assert np.array_equal(got[lo:hi], np.sort(a[lo:hi])), s


def test_segments_msd_graph_replays(mctx):
    """The launch sequence captured once and replayed with eager work between
    the replays: every replay sorts fresh keys."""
    import torch
    a, begin = _batch(0x6B, big=False)
    want = np.concatenate([a[: begin[0]]] + [np.sort(a[begin[s]:begin[s + 1]]) for s in range(begin.size - 1)]
                          + [a[begin[-1]:]])
    src = torch.from_numpy(a.view(np.int32).copy()).cuda()
    work = src.clone()
    b = torch.from_numpy(begin).cuda()
    N.set_option(N.HPCS_OPT_SEG_MSD, 16385, mctx)
    s = torch.cuda.Stream()
    call = lambda: N.check(N.lib().hpcs_dev_sort_segments_u32(mctx, P(work.data_ptr()), a.size, P(b.data_ptr()),
                                                              begin.size - 1, P(s.cuda_stream)))
    try:
        with torch.cuda.stream(s):
            call()  # sizes the scratch
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            call()
        for _ in range(3):
            work.copy_(src)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(work.cpu().numpy().view(np.uint32), want)
    finally:
        N.set_option(N.HPCS_OPT_SEG_MSD, 0, mctx)


def test_segments_msd_option_bounds(mctx):
    for bad in (1, 4096, 16384, 1 << 32, -1):
        with pytest.raises(N.HpcsError):
            N.set_option(N.HPCS_OPT_SEG_MSD, bad, mctx)
    N.set_option(N.HPCS_OPT_SEG_MSD, 0, mctx)
