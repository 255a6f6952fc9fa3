"""GPU parity: the memory-tight keys-only sort (HPCS_OPT_SCRATCH_LIMIT).

A limit one byte below the sort's full scratch makes the library sort each
half of the array in place and merge the halves with half the keys of
scratch (the GPU form of SortRadixInPlaceAdaptive's fallback to an in-place
sort on OutOfMemoryException, RadixSortMsd.cs:725-770).  Checked bit-exact
against the oracle for every key kind, on odd sizes (unequal halves), sizes
whose halves take the one-workgroup sort, and inputs that stress the merge
split (all equal, ascending, descending, two values, the halves' ranges
disjoint in either order).  The profile shows the merge ran.  A limit below
even the tight form's need fails with HPCS_E_BUDGET before any work (the
keys stay as they were), and so does a sort with values beyond the limit.
The last test takes almost all of the device's memory so that the full
sort's scratch allocation itself fails: the sort then takes the tight form.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p
FNS = {"u32": "hpcs_dev_sort_radix_u32", "u64": "hpcs_dev_sort_radix_u64", "f32": "hpcs_dev_sort_radix_f32",
       "f64": "hpcs_dev_sort_radix_f64", "i32": "hpcs_dev_sort_radix_i32", "i64": "hpcs_dev_sort_radix_i64"}
HOST_FNS = {"u32": "hpcs_sort_radix_u32", "u64": "hpcs_sort_radix_u64"}
DT = {"u32": np.uint32, "u64": np.uint64, "f32": np.float32, "f64": np.float64, "i32": np.int32, "i64": np.int64}
PLAN_DTYPE = {"u32": N.HPCS_DTYPE_U32, "i32": N.HPCS_DTYPE_U32, "f32": N.HPCS_DTYPE_F32,
              "u64": N.HPCS_DTYPE_U64, "i64": N.HPCS_DTYPE_U64, "f64": N.HPCS_DTYPE_F64}


@pytest.fixture()
def tctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = N.new_context(0)
    yield ctx
    torch.cuda.synchronize()
    N.destroy_context(ctx)


def _full_scratch(ctx, kind: str, n: int, with_values: bool = False) -> int:
    info = N.PlanInfo()
    N.check(N.lib().hpcs_plan_query(ctx, PLAN_DTYPE[kind], n, int(with_values), ctypes.byref(info)))
    return int(info.scratch_bytes)


def _expected(oracle, a: np.ndarray) -> np.ndarray:
    ref = a.copy()
    if a.dtype in (np.float32, np.float64):
        oracle.sort_radix_float(ref)
    elif a.dtype in (np.int32, np.int64):
        oracle.sort_radix_signed(ref)
    else:
        oracle.sort_radix_lsd(ref)
    return ref


def _keys(kind: str, n: int, seed: int, dist: str) -> np.ndarray:
    bits = 32 if kind in ("u32", "i32", "f32") else 64
    raw = (datagen.keys_u32(n, seed, "uniform").astype(np.uint64) if bits == 32
           else datagen.keys_u64(n, seed).astype(np.uint64))
    rng = np.random.default_rng(seed)
    h = (n + 1) // 2
    if dist == "equal":
        raw[:] = np.uint64(0x5A5A5A5A)
    elif dist == "ascending":
        raw = np.sort(raw)
    elif dist == "descending":
        raw = np.sort(raw)[::-1].copy()
    elif dist == "two":
        raw = np.where(rng.random(n) < 0.5, np.uint64(3), np.uint64(1 << (bits - 2)))
    elif dist == "low_high":     # first half's keys all below the second half's
        raw = np.sort(raw)
        raw[:h] = rng.permutation(raw[:h])
        raw[h:] = rng.permutation(raw[h:])
    elif dist == "high_low":     # first half's keys all above the second half's
        raw = np.sort(raw)[::-1].copy()
        raw[:h] = rng.permutation(raw[:h])
        raw[h:] = rng.permutation(raw[h:])
    if kind in ("f32", "f64"):
        raw &= np.uint64(0xBFFFFFFF if bits == 32 else 0xBFFFFFFFFFFFFFFF)  # finite: exponent never all ones
    a = raw.astype(np.uint32 if bits == 32 else np.uint64)
    return a.view(DT[kind])


def _dev_sort(ctx, a: np.ndarray, kind: str) -> tuple[int, np.ndarray]:
    import torch
    words = a.view(np.int32) if a.size else np.zeros(0, np.int32)
    buf = torch.from_numpy(words.copy()).cuda()
    rc = getattr(N.lib(), FNS[kind])(ctx, P(buf.data_ptr()), a.size, P(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, buf.cpu().numpy().view(a.dtype)


def _tight(ctx, kind: str, n: int) -> None:
    """Limit one byte below the full sort's scratch; profile on to see the merge."""
    N.set_option(N.HPCS_OPT_SCRATCH_LIMIT, _full_scratch(ctx, kind, n) - 1, ctx)
    N.profile_enable(True, ctx)


@pytest.mark.parametrize("kind", list(FNS))
@pytest.mark.parametrize("n", [3 * (1 << 20) + 7, (1 << 21) + 1, 40001, 30001])
def test_tight_sort_matches_oracle(oracle, tctx, kind, n):
    a = _keys(kind, n, 0x71C0 + n, "uniform")
    _tight(tctx, kind, n)
    rc, got = _dev_sort(tctx, a, kind)
    N.check(rc)
    assert "merge" in N.profile_read(tctx)
    assert np.array_equal(got.view(np.uint8), _expected(oracle, a).view(np.uint8))


@pytest.mark.parametrize("kind", ["u32", "u64", "f64", "i32"])
@pytest.mark.parametrize("dist", ["equal", "ascending", "descending", "two", "low_high", "high_low"])
def test_tight_sort_merge_splits(oracle, tctx, kind, dist):
    n = (1 << 20) + 3
    a = _keys(kind, n, 0xD157 + len(dist), dist)
    _tight(tctx, kind, n)
    rc, got = _dev_sort(tctx, a, kind)
    N.check(rc)
    assert "merge" in N.profile_read(tctx)
    assert np.array_equal(got.view(np.uint8), _expected(oracle, a).view(np.uint8))


@pytest.mark.parametrize("kind", list(HOST_FNS))
def test_tight_host_array_sort(oracle, tctx, kind):
    """The host-array entry point takes the same limit."""
    n = (1 << 21) + 5
    a = _keys(kind, n, 0xB057, "uniform")
    _tight(tctx, kind, n)
    b = np.ascontiguousarray(a.copy())
    N.check(getattr(N.lib(), HOST_FNS[kind])(tctx, b.ctypes.data_as(P), n))
    assert "merge" in N.profile_read(tctx)
    assert np.array_equal(b, _expected(oracle, a))


def test_tight_sort_graph_cache(oracle, tctx):
    """Three calls on one stream (direct, capture, replay) under the limit."""
    import torch
    n = (1 << 20) + 11
    a = _keys("u32", n, 0x6A, "uniform")
    N.set_option(N.HPCS_OPT_SCRATCH_LIMIT, _full_scratch(tctx, "u32", n) - 1, tctx)
    N.set_option(N.HPCS_OPT_GRAPHS, 1, tctx)
    s = torch.cuda.Stream()
    ref = _expected(oracle, a)
    src = torch.from_numpy(a.view(np.int32).copy()).cuda()
    buf = torch.empty_like(src)
    for _ in range(3):
        with torch.cuda.stream(s):
            buf.copy_(src)
            N.check(N.lib().hpcs_dev_sort_radix_u32(tctx, P(buf.data_ptr()), n, P(s.cuda_stream)))
        s.synchronize()
        assert np.array_equal(buf.cpu().numpy().view(np.uint32), ref)
    torch.cuda.synchronize()


@pytest.mark.parametrize("kind", ["u32", "u64"])
def test_limit_below_tight_need_is_budget_error(tctx, kind):
    n = 1 << 20
    a = _keys(kind, n, 0xBAD, "uniform")
    N.set_option(N.HPCS_OPT_SCRATCH_LIMIT, 4096, tctx)
    rc, got = _dev_sort(tctx, a, kind)
    assert rc == N.HPCS_E_BUDGET
    assert np.array_equal(got, a)            # no work before the error
    N.set_option(N.HPCS_OPT_SCRATCH_LIMIT, 0, tctx)
    rc, got = _dev_sort(tctx, a, kind)     # no limit: the full sort
    N.check(rc)
    assert np.array_equal(got, np.sort(a))


def test_pairs_beyond_limit_is_budget_error(tctx):
    import torch
    n = 1 << 20
    keys = torch.from_numpy(datagen.keys_u32(n, 0xFA, "uniform").view(np.int32).copy()).cuda()
    vals = torch.arange(n, dtype=torch.int32, device="cuda")
    k0 = keys.clone()
    N.set_option(N.HPCS_OPT_SCRATCH_LIMIT, _full_scratch(tctx, "u32", n, True) - 1, tctx)
    st = P(torch.cuda.current_stream().cuda_stream)
    rc = N.lib().hpcs_dev_sort_radix_pairs_u32_u32(tctx, P(keys.data_ptr()), P(vals.data_ptr()), n, st)
    torch.cuda.synchronize()
    assert rc == N.HPCS_E_BUDGET
    assert torch.equal(keys, k0)
    N.set_option(N.HPCS_OPT_SCRATCH_LIMIT, _full_scratch(tctx, "u32", n, True), tctx)
    N.check(N.lib().hpcs_dev_sort_radix_pairs_u32_u32(tctx, P(keys.data_ptr()), P(vals.data_ptr()), n, st))
    torch.cuda.synchronize()
    order = np.argsort(k0.cpu().numpy().view(np.uint32), kind="stable")
    assert np.array_equal(vals.cpu().numpy(), order.astype(np.int32))


def test_out_of_memory_takes_tight_form(oracle, tctx):
    """Hold all but ~3/4 of the full sort's scratch of device memory: the
    full sort's scratch allocation fails with out-of-memory and the sort
    takes the tight form (about half the scratch), as the reference's catch
    does."""
    import torch
    n = 1 << 27
    a = _keys("u32", n, 0x00A, "uniform")
    ref = _expected(oracle, a)
    buf = torch.from_numpy(a.view(np.int32).copy()).cuda()
    full = _full_scratch(tctx, "u32", n)
    leave = (full * 3) // 4                    # above the tight need (~half), below the full need
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    holds, seen = [], []
    try:
        for _ in range(4):                     # the allocator rounds: close in on the target
            free, _ = torch.cuda.mem_get_info()
            seen.append(free)
            if free <= full - (16 << 20):
                break
            try:
                holds.append(torch.empty(free - leave, dtype=torch.uint8, device="cuda"))
            except torch.OutOfMemoryError:
                holds.append(torch.empty((free - leave) // 2, dtype=torch.uint8, device="cuda"))
        free, _ = torch.cuda.mem_get_info()
        seen.append(free)
        if not (full // 2 + (64 << 20) < free < full):
            pytest.skip(f"free device memory {seen} not between the tight and full needs (full {full})")
        N.profile_enable(True, tctx)
        N.check(N.lib().hpcs_dev_sort_radix_u32(tctx, P(buf.data_ptr()), n,
                                                P(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        assert "merge" in N.profile_read(tctx)
    finally:
        del holds
        torch.cuda.empty_cache()
    assert np.array_equal(buf.cpu().numpy().view(np.uint32), ref)
