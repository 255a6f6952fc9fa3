"""GPU parity: the MSD-first form of the uint32 keys-only sort
(HPCS_OPT_PASS_ALGO 6).

Bit-exact against the oracle's SortRadixLsd restatement (and np.sort at the
larger sizes) on inputs chosen to reach every branch the device plans:
- uniform 32-bit keys (top-byte split counted by the upsweep, then byte 2,
  then every (byte 3, byte 2) bucket in LDS);
- dense keys (0 .. n-1, and uniform keys below 2^24 / 2^25): the split digits
  move down to the top varying bits and a counting kernel recounts them;
- keys whose bits from 23 up never vary (low 8 bits only, constant, below
  2^23): the plan falls back to the plain LSD passes inside the same launches;
- small arrays, whose top-byte buckets hold at most 4096 keys and are sorted
  whole by the batched segments' LDS kernel;
- one top byte holding 90% of the keys, few distinct values, keys whose low
  bits are all equal: buckets whose 12-bit bins outgrow the rank loop (the
  1024-thread overflow sort) and buckets beyond 16384 keys (two segmented
  overflow passes);
- float and signed keys, which take the reduce-then-scan passes with the same
  table row.
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
def hctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = N.new_context(0)
    N.set_option(N.HPCS_OPT_PASS_ALGO, 6, ctx)
    yield ctx
    torch.cuda.synchronize()
    N.destroy_context(ctx)


def _keys(n: int, seed: int, dist: str) -> np.ndarray:
    r = datagen.keys_u32(n, seed, "uniform")
    if dist == "uniform":
        return r
    if dist == "dense":
        return np.random.default_rng(seed).permutation(n).astype(np.uint32)
    if dist == "ascending":
        return np.arange(n, dtype=np.uint32)
    if dist == "below2^24":
        return r & np.uint32(0xFFFFFF)
    if dist == "below2^25":
        return r & np.uint32(0x1FFFFFF)
    if dist == "below2^23":
        return r & np.uint32(0x7FFFFF)
    if dist == "low8":
        return (r & np.uint32(0xFF)) | np.uint32(0x12345600)
    if dist == "constant":
        return np.full(n, 0xDEADBEEF, np.uint32)
    if dist == "skew90":
        return np.where(r % 10 < 9, (r & np.uint32(0xFFFFFF)) | np.uint32(0x5A000000), r).astype(np.uint32)
    if dist == "fewvals":  # 1000 distinct values spread over the range: every bucket one bin
        return ((r % 1000) * np.uint32(4294967)).astype(np.uint32)
    if dist == "lowzero":  # low 12 bits zero: 16 keys per 12-bit bin
        return r & np.uint32(0xFFFFF000)
    if dist == "top16":  # 16 top-16-bit values: buckets far beyond 16384 keys
        return ((r % 16) << np.uint32(16)) | (r & np.uint32(0xFFFF))
    if dist == "top2k":  # 2048 top-16-bit values: buckets of ~n/2048 keys (medium for n = 2^24)
        return (((r >> np.uint32(5)) % 2048) << np.uint32(16)) | (r & np.uint32(0xFFFF))
    raise ValueError(dist)


def _sort(ctx, a: np.ndarray, fn: str = "hpcs_dev_sort_radix_u32") -> np.ndarray:
    import torch
    buf = torch.from_numpy(a.view(np.int32).copy()).cuda()
    N.check(getattr(N.lib(), fn)(ctx, P(buf.data_ptr()), a.size, P(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return buf.cpu().numpy().view(a.dtype)


DISTS = ["uniform", "dense", "ascending", "below2^24", "below2^25", "below2^23", "low8", "constant", "skew90",
         "fewvals", "lowzero", "top16", "top2k"]


@pytest.mark.parametrize("dist", DISTS)
@pytest.mark.parametrize("n", [16385, 100003, (1 << 20) + 77, 3 * (1 << 20) + 5])
def test_hybrid_matches_oracle(oracle, hctx, n, dist):
    a = _keys(n, 0x4B1D + n, dist)
    want = a.copy()
    oracle.sort_radix_lsd(want)
    got = _sort(hctx, a)
    assert np.array_equal(got, want), f"n={n} {dist}"


@pytest.mark.parametrize("dist", ["uniform", "dense", "skew90", "top16", "top2k", "lowzero"])
def test_hybrid_2pow24(hctx, dist):
    n = 1 << 24
    a = _keys(n, 0x24 + len(dist), dist)
    assert np.array_equal(_sort(hctx, a), np.sort(a))


def test_hybrid_profile_shows_stages(hctx):
    """Uniform keys and keys below 2^23 (the LSD passes through the same
    stage slots) run one launch sequence: one global pass slot, the
    segmented pass, the bucket sort and the overflow passes."""
    n = 1 << 22
    for dist in ("uniform", "below2^23", "low8"):
        a = _keys(n, 0x7A6 + len(dist), dist)
        N.profile_enable(True, hctx)
        got = _sort(hctx, a)
        prof = N.profile_read(hctx)
        N.profile_enable(False, hctx)
        assert np.array_equal(got, np.sort(a)), dist
        assert "hyb_local" in prof and "seg_scatter" in prof, prof
        assert prof["scatter"][0] == 1 and prof["tile_counts"][0] == 1, prof  # the one global pass slot
        assert prof["seg_scatter"][0] == 3, prof  # the segmented pass and two overflow passes


def test_hybrid_float_and_signed_take_the_passes(oracle, hctx):
    """Algo 6 in a 4-byte row also serves float and signed keys, with the
    reduce-then-scan passes."""
    n = (1 << 20) + 3
    with np.errstate(over="ignore"):  # +-1e300 become +-inf in float32
        x = datagen.doubles_special(n, 11).astype(np.float32)
    want = x.copy()
    oracle.sort_radix_float(want)
    got = _sort(hctx, x, "hpcs_dev_sort_radix_f32")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    i = datagen.keys_u32(n, 12, "uniform").view(np.int32)
    assert np.array_equal(_sort(hctx, i, "hpcs_dev_sort_radix_i32"), np.sort(i))


def test_hybrid_graph_replays(hctx):
    """The sequence captured once in a HIP graph and replayed with eager work
    between the replays: every replay sorts fresh keys correctly."""
    import torch
    n = (1 << 22) + 9
    a = _keys(n, 0x6A, "uniform")
    src = torch.from_numpy(a.view(np.int32).copy()).cuda()
    work = src.clone()
    s = torch.cuda.Stream()
    call = lambda: N.check(N.lib().hpcs_dev_sort_radix_u32(hctx, P(work.data_ptr()), n, P(s.cuda_stream)))
    with torch.cuda.stream(s):
        call()  # sizes the scratch
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call()
    want = np.sort(a)
    for _ in range(3):
        work.copy_(src)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(work.cpu().numpy().view(np.uint32), want)
