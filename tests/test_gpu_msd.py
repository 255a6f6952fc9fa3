"""GPU parity: the MSD split + one workgroup per bucket (HPCS_OPT_PASS_ALGO 4).

Bit-exact against numpy for every key kind, sizes from 2 keys to the path's
largest (2^22 4-byte, 2^21 8-byte keys), and inputs that stress the split:
keys using part of the range, one bucket holding most keys (the streaming
bucket passes), two distinct values, all equal, ascending, descending, and
float specials.  Stability is checked with a key array whose ties carry
their input position in the low bits, which the sort must not reorder.
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
DT = {"u32": np.uint32, "u64": np.uint64, "f32": np.float32, "f64": np.float64, "i32": np.int32, "i64": np.int64}


@pytest.fixture(scope="module")
def mctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = N.new_context(0)
    N.set_option(N.HPCS_OPT_PASS_ALGO, 4, ctx)
    yield ctx
    torch.cuda.synchronize()
    N.destroy_context(ctx)


def _expected(a: np.ndarray) -> np.ndarray:
    if a.dtype in (np.float32, np.float64):
        u = a.view(np.uint32 if a.dtype == np.float32 else np.uint64)
        sign = u.dtype.type(1) << u.dtype.type(8 * a.itemsize - 1)
        t = np.where(u & sign, ~u, u ^ sign)
        return a[np.argsort(t, kind="stable")]
    return np.sort(a, kind="stable")


def _sort(ctx, a: np.ndarray) -> np.ndarray:
    import torch
    kind = {v: k for k, v in DT.items()}[a.dtype.type]
    words = a.view(np.int32) if a.size else np.zeros(0, np.int32)
    buf = torch.from_numpy(words.copy()).cuda()
    N.check(getattr(N.lib(), FNS[kind])(ctx, P(buf.data_ptr()), a.size, P(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    return buf.cpu().numpy().view(a.dtype)


def _keys(kind: str, n: int, seed: int, dist: str) -> np.ndarray:
    bits = 32 if kind in ("u32", "i32", "f32") else 64
    raw = datagen.keys_u32(n, seed, "uniform").astype(np.uint64) if bits == 32 else datagen.keys_u64(n, seed)
    raw = raw.astype(np.uint64)
    rng = np.random.default_rng(seed)
    if dist == "partial":      # keys below 2^(bits - 11): the split digit is not the top byte
        raw &= np.uint64((1 << (bits - 11)) - 1)
    elif dist == "skewed":     # 90% of keys in one bucket: the streaming bucket passes
        hot = rng.random(n) < 0.9
        raw = np.where(hot, raw & np.uint64(0xFFFF), raw)
    elif dist == "two":
        raw = np.where(rng.random(n) < 0.5, np.uint64(3), np.uint64(1 << (bits - 2)))
    elif dist == "equal":
        raw[:] = np.uint64(0x5A5A5A5A)
    elif dist == "ascending":
        raw = np.arange(n, dtype=np.uint64) * np.uint64(2654435761 if bits == 32 else 11400714819323198485)
        raw = np.sort(raw & np.uint64((1 << bits) - 1))
    elif dist == "descending":
        raw = np.sort(raw)[::-1].copy()
    if kind in ("f32", "f64"):
        mask = 0xBFFFFFFF if bits == 32 else 0xBFFFFFFFFFFFFFFF  # finite: exponent never all ones
        raw &= np.uint64(mask)
    a = raw.astype(np.uint32 if bits == 32 else np.uint64)
    if kind in ("f32", "f64") and n >= 8:
        f = a.view(DT[kind])
        f[:4] = [0.0, -0.0, np.finfo(DT[kind]).tiny / 4, -np.finfo(DT[kind]).tiny / 2]
    return a.view(DT[kind])


@pytest.mark.parametrize("kind", ["u32", "u64", "f32", "f64", "i32", "i64"])
@pytest.mark.parametrize("n", [2, 3, 1000, 4096, 65543, 1 << 20])
def test_uniform_sizes(mctx, kind, n):
    a = _keys(kind, n, 0x7100 + n, "uniform")
    assert np.array_equal(_sort(mctx, a).view(np.uint8), _expected(a).view(np.uint8))


@pytest.mark.parametrize("kind,n", [("u32", 1 << 22), ("u32", (1 << 21) + 5), ("u64", 1 << 21), ("f64", 1 << 21)])
def test_largest_sizes(mctx, kind, n):
    a = _keys(kind, n, 0x7200 + n, "uniform")
    assert np.array_equal(_sort(mctx, a).view(np.uint8), _expected(a).view(np.uint8))


@pytest.mark.parametrize("kind", ["u32", "u64", "f64", "i32"])
@pytest.mark.parametrize("dist", ["partial", "skewed", "two", "equal", "ascending", "descending"])
def test_distributions(mctx, kind, dist):
    n = 1 << 18
    a = _keys(kind, n, 0x7300, dist)
    assert np.array_equal(_sort(mctx, a).view(np.uint8), _expected(a).view(np.uint8))


def test_stable_ties(mctx):
    """Keys whose top 16 bits repeat and whose low 16 bits hold the input
    position: a stable sort by the full key equals a stable sort by the top
    bits alone, which the sort's order must reproduce exactly."""
    n = 1 << 16
    top = datagen.keys_u32(n, 0x7400, "uniform") & np.uint32(0xFF000000)
    a = top | np.arange(n, dtype=np.uint32)
    got = _sort(mctx, a)
    assert np.array_equal(got, a[np.argsort(top, kind="stable")])


def test_profile_names_msd_kernels(mctx):
    n = 1 << 20
    a = _keys("u32", n, 0x7500, "uniform")
    N.profile_enable(True, mctx)
    try:
        _sort(mctx, a)
        prof = N.profile_read(mctx)
    finally:
        N.profile_enable(False, mctx)
    assert {"msd_or_and", "msd_hist", "msd_split", "msd_buckets"} <= set(prof)
    assert "onesweep" not in prof and "scatter" not in prof


def test_plan_query_reports_the_split(mctx):
    info = N.PlanInfo()
    n = 1 << 20
    N.check(N.lib().hpcs_plan_query(mctx, N.HPCS_DTYPE_U32, n, 0, ctypes.byref(info)))
    assert (info.digit_bits, info.passes, info.ideal_bytes) == (8, 2, 6 * 4 * n)
    # key-value sorts and arrays beyond the split's range take the by-size plan
    N.check(N.lib().hpcs_plan_query(mctx, N.HPCS_DTYPE_U32, n, 1, ctypes.byref(info)))
    assert info.passes == 4
    N.check(N.lib().hpcs_plan_query(mctx, N.HPCS_DTYPE_U64, 1 << 22, 0, ctypes.byref(info)))
    assert info.passes == 8


@pytest.mark.parametrize("kind", ["u32", "u64"])
@pytest.mark.parametrize("dist", ["uniform", "skewed", "partial"])
def test_split_forced_lookback_fallback(oracle, kind, dist):
    """The MSD split's look-back pass with every wait resolved by counting
    the earlier tile locally (fallback threshold 0): still the oracle's
    output, on spread and on clustered keys."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, 4, ctx)
        N.set_option(N.HPCS_OPT_FALLBACK_ROUNDS, 0, ctx)
        n = (1 << 20) + 4099 if kind == "u32" else (1 << 19) + 77
        a = _keys(kind, n, 31 + n, dist)
        ref = a.copy()
        oracle.sort_radix_lsd(ref)
        assert np.array_equal(_sort(ctx, a), ref), (kind, dist)
    finally:
        N.destroy_context(ctx)
