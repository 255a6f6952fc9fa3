"""GPU: the context's graph cache (HPCS_OPT_GRAPHS) for device-resident sorts.

A repeated call (same entry point, pointers, size and stream) is replayed
from a captured HIP graph.  Every replay runs on fresh input and is checked
bit-exact against numpy (keys) or a stable argsort (pairs); the cache must
recapture when the context's scratch moves, clear on an option change, stay
out of the way under a caller's own capture and on the null stream.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p


@pytest.fixture()
def gctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = N.new_context(0)
    N.set_option(N.HPCS_OPT_GRAPHS, 1, ctx)
    yield ctx
    torch.cuda.synchronize()
    N.destroy_context(ctx)


def _keys(kind: str, n: int, seed: int) -> np.ndarray:
    if kind in ("u32", "i32", "f32"):
        a = datagen.keys_u32(n, seed, "uniform")
        if kind == "f32":  # finite bit patterns: exponent 255 cleared
            a = a & np.uint32(0xBFFFFFFF)
        return a.view({"u32": np.uint32, "i32": np.int32, "f32": np.float32}[kind])
    a = datagen.keys_u64(n, seed)
    return a.view({"u64": np.uint64, "i64": np.int64, "f64": np.float64}[kind])


def _expected(kind: str, a: np.ndarray) -> np.ndarray:
    if kind in ("f32", "f64"):
        # ascending with -0.0 before +0.0: sort by the radix key transform
        u = a.view(np.uint32 if kind == "f32" else np.uint64)
        sign = np.uint32(1 << 31) if kind == "f32" else np.uint64(1 << 63)
        t = np.where(u & sign, ~u, u ^ sign)
        return a[np.argsort(t, kind="stable")]
    return np.sort(a)


FNS = {"u32": "hpcs_dev_sort_radix_u32", "u64": "hpcs_dev_sort_radix_u64", "f32": "hpcs_dev_sort_radix_f32",
       "f64": "hpcs_dev_sort_radix_f64", "i32": "hpcs_dev_sort_radix_i32", "i64": "hpcs_dev_sort_radix_i64"}


# sizes cover each launch sequence: one workgroup, onesweep passes, the lean
# and the full reduce-then-scan forms
@pytest.mark.parametrize("kind,n", [("u32", 4000), ("u32", 1 << 16), ("u32", 1 << 20), ("u32", (1 << 22) + 3),
                                    ("u64", 1 << 18), ("f32", 100_003), ("f64", 1 << 21), ("i32", 70_000),
                                    ("i64", 9000)])
def test_replays_match_numpy(gctx, kind, n):
    import torch
    s = torch.cuda.Stream()
    f = getattr(N.lib(), FNS[kind])
    kb = 4 if kind in ("u32", "i32", "f32") else 8
    buf = torch.empty(n * kb // 4, dtype=torch.int32, device="cuda")
    e0, c0, r0 = N.graph_cache_stats(gctx)
    for rep in range(5):
        a = _keys(kind, n, 0x6A00 + rep)
        buf.copy_(torch.from_numpy(a.view(np.int32)))
        torch.cuda.synchronize()
        N.check(f(gctx, P(buf.data_ptr()), n, P(s.cuda_stream)))
        s.synchronize()
        got = buf.cpu().numpy().view(a.dtype)
        assert np.array_equal(got.view(np.uint8), _expected(kind, a).view(np.uint8)), f"call {rep}"
    e1, c1, r1 = N.graph_cache_stats(gctx)
    assert c1 - c0 == 1 and r1 - r0 == 4 and e1 >= 1


def test_pairs_replays_are_stable(gctx):
    import torch
    s = torch.cuda.Stream()
    n = 300_000
    k = torch.empty(2 * n, dtype=torch.int32, device="cuda")
    v = torch.empty(n, dtype=torch.int32, device="cuda")
    for rep in range(4):
        keys = datagen.keys_u64(n, 0x6B00 + rep) & np.uint64(0xFFFF)  # many ties
        vals = np.arange(n, dtype=np.uint32)
        k.copy_(torch.from_numpy(keys.view(np.int32)))
        v.copy_(torch.from_numpy(vals.view(np.int32)))
        torch.cuda.synchronize()
        N.check(N.lib().hpcs_dev_sort_radix_pairs_u64_u32(gctx, P(k.data_ptr()), P(v.data_ptr()), n,
                                                          P(s.cuda_stream)))
        s.synchronize()
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k.cpu().numpy().view(np.uint64), keys[order])
        assert np.array_equal(v.cpu().numpy().view(np.uint32), vals[order])
    assert N.graph_cache_stats(gctx)[1:] == (1, 3)


def test_scratch_growth_recaptures(gctx):
    import torch
    s = torch.cuda.Stream()
    f = N.lib().hpcs_dev_sort_radix_u32
    small, big = 1 << 16, 1 << 22
    bs = torch.empty(small, dtype=torch.int32, device="cuda")
    bb = torch.empty(big, dtype=torch.int32, device="cuda")

    def run(buf, n, seed):
        a = datagen.keys_u32(n, seed, "uniform")
        buf.copy_(torch.from_numpy(a.view(np.int32)))
        torch.cuda.synchronize()
        N.check(f(gctx, P(buf.data_ptr()), n, P(s.cuda_stream)))
        s.synchronize()
        assert np.array_equal(buf.cpu().numpy().view(np.uint32), np.sort(a))

    run(bs, small, 1)
    run(bs, small, 2)          # captured
    run(bs, small, 3)          # replayed
    run(bb, big, 4)            # scratch grows: the small call's graph addresses the old buffer
    run(bs, small, 5)          # runs directly again (buffers moved)
    run(bs, small, 6)          # recaptured
    run(bs, small, 7)          # replayed
    assert N.graph_cache_stats(gctx)[1:] == (2, 4)


def test_option_change_clears_and_user_capture_is_respected(gctx):
    import torch
    s = torch.cuda.Stream()
    f = N.lib().hpcs_dev_sort_radix_u32
    n = 1 << 18
    buf = torch.empty(n, dtype=torch.int32, device="cuda")
    src = torch.from_numpy(datagen.keys_u32(n, 9, "uniform").view(np.int32)).cuda()
    want = np.sort(src.cpu().numpy().view(np.uint32))
    for _ in range(2):
        buf.copy_(src)
        torch.cuda.synchronize()
        N.check(f(gctx, P(buf.data_ptr()), n, P(s.cuda_stream)))
        s.synchronize()
    assert N.graph_cache_stats(gctx)[0] == 1
    N.set_option(N.HPCS_OPT_RANK, 0, gctx)
    assert N.graph_cache_stats(gctx)[0] == 0
    # the caller captures: the library launches plainly into the caller's graph
    before = N.graph_cache_stats(gctx)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        buf.copy_(src)
        N.check(f(gctx, P(buf.data_ptr()), n, P(s.cuda_stream)))
    assert N.graph_cache_stats(gctx) == before
    for _ in range(3):
        buf.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(buf.cpu().numpy().view(np.uint32), want)
    del g
    # the null stream never goes through the cache
    for _ in range(3):
        buf.copy_(src)
        torch.cuda.synchronize()
        N.check(f(gctx, buf.data_ptr(), n, None))
        torch.cuda.synchronize()
        assert np.array_equal(buf.cpu().numpy().view(np.uint32), want)
    assert N.graph_cache_stats(gctx) == before


def test_segments_replays(gctx):
    import torch
    s = torch.cuda.Stream()
    nseg = 300
    lens = datagen.segment_lengths(nseg, 0x6C00, 1, 1 << 14)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(begin[-1])
    buf = torch.empty(total, dtype=torch.int32, device="cuda")
    d_begin = torch.from_numpy(begin).cuda()
    for rep in range(4):
        a = datagen.keys_u32(total, 0x6C10 + rep, "uniform")
        buf.copy_(torch.from_numpy(a.view(np.int32)))
        torch.cuda.synchronize()
        N.check(N.lib().hpcs_dev_sort_segments_u32(gctx, P(buf.data_ptr()), total, P(d_begin.data_ptr()), nseg,
                                                   P(s.cuda_stream)))
        s.synchronize()
        got = buf.cpu().numpy().view(np.uint32)
        for i in range(nseg):
            assert np.array_equal(got[begin[i]:begin[i + 1]], np.sort(a[begin[i]:begin[i + 1]])), (rep, i)
    assert N.graph_cache_stats(gctx)[1:] == (1, 3)


def test_segments_replay_follows_new_boundaries(gctx):
    """The key holds the offsets pointer, not its contents: the segment
    boundaries are read on the device, so a replay with new boundaries in the
    same buffer (same total and count) must sort the new segments."""
    import torch
    s = torch.cuda.Stream()
    nseg, total = 200, 1 << 20
    buf = torch.empty(total, dtype=torch.int32, device="cuda")
    d_begin = torch.empty(nseg + 1, dtype=torch.int64, device="cuda")
    rng = np.random.default_rng(0x6D00)
    for rep in range(4):
        cuts = np.sort(rng.choice(np.arange(1, total), size=nseg - 1, replace=False))
        begin = np.concatenate([[0], cuts, [total]]).astype(np.int64)
        if rep == 3:  # one large segment and many empty ones
            begin = np.concatenate([np.zeros(nseg // 2, np.int64), np.full(nseg - nseg // 2 + 1, total, np.int64)])
        a = datagen.keys_u32(total, 0x6D10 + rep, "uniform")
        buf.copy_(torch.from_numpy(a.view(np.int32)))
        d_begin.copy_(torch.from_numpy(begin))
        torch.cuda.synchronize()
        N.check(N.lib().hpcs_dev_sort_segments_u32(gctx, P(buf.data_ptr()), total, P(d_begin.data_ptr()), nseg,
                                                   P(s.cuda_stream)))
        s.synchronize()
        got = buf.cpu().numpy().view(np.uint32)
        for i in range(nseg):
            assert np.array_equal(got[begin[i]:begin[i + 1]], np.sort(a[begin[i]:begin[i + 1]])), (rep, i)
    assert N.graph_cache_stats(gctx)[1:] == (1, 3)
