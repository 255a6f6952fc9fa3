"""GPU parity of the radix top-k against the oracle's SelectTopK restatement
(u32) and the closed form (the k largest = the tail of np.sort) for every key
type.  Bit-exact: keys are compared as raw bits, floats in the radix order
(-0.0 below +0.0)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N
from hpcsharp_amd.algorithm import Algorithm, ArgumentNullException, ArgumentOutOfRangeException
from tests._fixtures import float_order

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    N.context(0)
    yield


def _expected_desc(x: np.ndarray, k: int) -> np.ndarray:
    if x.dtype.kind == "f":
        return float_order(x)[::-1][:k]
    return np.sort(x)[::-1][:k]


def _bits(x: np.ndarray) -> np.ndarray:
    return x.view(np.uint32 if x.itemsize == 4 else np.uint64)


def _keys(dtype, n, seed, dist="uniform"):
    if dtype == np.uint32:
        return datagen.keys_u32(n, seed, dist)
    if dtype == np.uint64:
        k = datagen.keys_u64(n, seed)
        if dist == "low8":
            k = (k & np.uint64(0xFF)) | np.uint64(0xABCDEF0123456700)
        elif dist == "sorted":
            k = np.sort(k)
        elif dist == "constant":
            k = np.full(n, 0x0123456789ABCDEF, dtype=np.uint64)
        return k
    if dtype == np.float32:
        x = datagen.floats_finite(n, seed)
        if dist == "constant":
            x = np.full(n, -2.5, dtype=np.float32)
        return x
    x = datagen.doubles_special(n, seed)
    if dist == "constant":
        x = np.full(n, 1e300, dtype=np.float64)
    return x


@pytest.mark.parametrize("n,k", [(1000, 10), (65537, 1), (65537, 4096), (1 << 20, 10000), (300001, 16384),
                                 (300001, 16385), (200003, 200003), (5, 5), (1, 1), (12345, 0),
                                 (16384, 16384), (9000, 4500), (2049, 1025), (20000, 1024), (256, 256),
                                 (257, 200), (100000, 256), (100000, 255), (70000, 1)])
def test_topk_u32_vs_oracle(oracle, n, k):
    a = datagen.keys_u32(n, n ^ k)
    got = Algorithm.TopK(a, k)
    ref = a.copy()
    if 0 < k < n:
        oracle.select_topk_u32(ref, n - k)    # a[n-k:] = the k largest
    top = np.sort(ref[n - k:])[::-1]
    assert np.array_equal(got, top)
    assert np.array_equal(got, _expected_desc(a, k))


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.float32, np.float64])
@pytest.mark.parametrize("dist", ["uniform", "low8", "sorted", "constant"])
@pytest.mark.parametrize("n,k", [(100003, 777), (70001, 20000)])
def test_topk_types_and_distributions(dtype, dist, n, k):
    if dtype in (np.float32, np.float64) and dist in ("low8", "sorted"):
        pytest.skip("distribution defined for unsigned keys")
    x = _keys(dtype, n, 31 + k, dist)
    orig = x.copy()
    got = Algorithm.TopK(x, k)
    assert np.array_equal(_bits(x), _bits(orig))          # input untouched
    assert np.array_equal(_bits(got), _bits(_expected_desc(orig, k)))


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.float32, np.float64])
@pytest.mark.parametrize("dist", ["uniform", "low8", "sorted", "constant"])
def test_topk_sampled_filter(dtype, dist):
    """Inputs of at least 2^20 keys with k <= n / 64 are sampled and their top
    level-0 bins kept in one read before the descent: every key kind and
    distribution (constant keys: every key kept), k from 1 to the limit."""
    if dtype in (np.float32, np.float64) and dist in ("low8", "sorted"):
        pytest.skip("distribution defined for unsigned keys")
    n = (1 << 20) + 77
    x = _keys(dtype, n, 515, dist)
    for k in (1, 1000, n // 64):
        got = Algorithm.TopK(x, k)
        assert np.array_equal(_bits(got), _bits(_expected_desc(x, k))), k


def test_topk_sampled_filter_falls_back():
    """A sample that overstates the keys (the sampled positions hold the only
    large keys, and k = n / 64 exceeds the 2^16 sampled keys) keeps fewer
    than k keys: level 0 then reads the whole input.  Same result."""
    n = 1 << 23
    stride = n >> 16
    a = datagen.keys_u32(n, 77) & np.uint32(0x000FFFFF)
    pos = np.arange(1 << 16, dtype=np.int64) * stride + stride // 2
    a[pos] |= np.uint32(0xFFF00000)
    for k in (n // 64, n // 100):
        assert np.array_equal(Algorithm.TopK(a, k), _expected_desc(a, k)), k


@pytest.mark.parametrize("mask", [0xF, 0x3FF, 0xFFFF0000, 0x80000001])
def test_topk_ties(mask):
    """Narrow keys: the threshold key has many copies, some needed, some not."""
    n = 250007
    a = datagen.keys_u32(n, mask) & np.uint32(mask)
    for k in (1, 100, 12345, 100000, n - 1):
        assert np.array_equal(Algorithm.TopK(a, k), _expected_desc(a, k)), k


def test_topk_float_specials():
    x = np.array([0.0, -0.0, 1.0, -1.0, 5e-324, -5e-324, 1e300, -1e300, 0.0, -0.0], dtype=np.float64)
    got = Algorithm.TopK(x, 6)
    assert _bits(got).tolist() == _bits(np.array([1e300, 1.0, 5e-324, 0.0, 0.0, -0.0])).tolist()
    with np.errstate(over="ignore"):
        y = x.astype(np.float32)                         # +-1e300 become +-inf
    got = Algorithm.TopK(y, 10)
    assert np.array_equal(_bits(got), _bits(float_order(y)[::-1]))


@pytest.mark.parametrize("n,k", [(1 << 16, 0), (1 << 16, 1), (1 << 16, 40000), ((1 << 20) + 5, 777777),
                                 (1000, 999), (3, 1)])
def test_select_topk_in_place_vs_oracle(oracle, n, k):
    """SelectTopK(a, k): a[k:] = the n-k largest ascending; a stays a
    permutation; a[k] equals the oracle's k-th element."""
    a = datagen.keys_u32(n, 5 + k)
    ref = a.copy()
    oracle.select_topk_u32(ref, k)
    orig = a.copy()
    Algorithm.SelectTopK(a, k)
    assert a[k] == ref[k]
    assert np.array_equal(a[k:], np.sort(ref[k:]))
    assert np.array_equal(np.sort(a), np.sort(orig))


def test_select_topk_sub_range(oracle):
    n = 100000
    a = datagen.keys_u32(n, 99) & np.uint32(0xFFF)
    orig = a.copy()
    start, length, k = 1234, 50000, 20000
    Algorithm.SelectTopK(a, start, length, k)
    assert np.array_equal(a[:start], orig[:start])
    assert np.array_equal(a[start + length:], orig[start + length:])
    seg = orig[start:start + length].copy()
    oracle.select_topk_u32(seg, k - start)
    assert np.array_equal(a[k:start + length], np.sort(seg[k - start:]))
    assert np.array_equal(np.sort(a[start:start + length]), np.sort(orig[start:start + length]))


def test_topk_argument_errors():
    a = datagen.keys_u32(100, 1)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.TopK(a, 101)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.TopK(a, -1)
    with pytest.raises(ArgumentNullException):
        Algorithm.SelectTopK(None, 3)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.SelectTopK(a, 100)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.SelectTopK(a, 10, 50, 5)
    rc = N.lib().hpcs_topk_u32(N.context(), None, 10, 1, None)
    assert rc == N.HPCS_E_INVALID_ARG
    assert Algorithm.TopK(a, 0).size == 0


_DEV = {np.uint32: "hpcs_dev_topk_u32", np.uint64: "hpcs_dev_topk_u64",
        np.float32: "hpcs_dev_topk_f32", np.float64: "hpcs_dev_topk_f64"}


def _to_dev(x: np.ndarray):
    import torch
    v = x.view(np.int32 if x.itemsize == 4 else np.int64)
    return torch.from_numpy(v.copy()).cuda()


def _from_dev(t, dtype):
    return t.cpu().numpy().view(dtype)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.float32, np.float64])
def test_dev_topk_on_stream(dtype):
    import torch
    n, k = (1 << 21) + 19, 5000
    x = _keys(dtype, n, 4242)
    t = _to_dev(x)
    out = torch.empty(k, dtype=t.dtype, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        N.check(getattr(N.lib(), _DEV[dtype])(N.context(), ctypes.c_void_p(t.data_ptr()), n, k,
                                              ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
    s.synchronize()
    assert np.array_equal(_bits(_from_dev(out, dtype)), _bits(_expected_desc(x, k)))


def test_dev_topk_graph_replays_match_eager():
    """Captured once in a HIP graph: every replay, with eager work between
    replays and new input contents, equals the eager result."""
    import torch
    n, k = 1 << 20, 3000
    inputs = [datagen.keys_u32(n, 900 + r) & np.uint32(0xFFFFF) for r in range(3)]
    src = [_to_dev(a) for a in inputs]
    work = src[0].clone()
    out = torch.empty(k, dtype=torch.int32, device="cuda")
    ctx = N.context()
    s = torch.cuda.Stream()

    def call():
        N.check(N.lib().hpcs_dev_topk_u32(ctx, ctypes.c_void_p(work.data_ptr()), n, k,
                                          ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(s.cuda_stream)))
    with torch.cuda.stream(s):
        call()                                                     # warm-up sizes scratch
    s.synchronize()
    assert np.array_equal(_from_dev(out, np.uint32), _expected_desc(inputs[0], k))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call()
    for rep in range(3):
        work.copy_(src[rep])
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(_from_dev(out, np.uint32), _expected_desc(inputs[rep], k)), rep
        _ = (src[0][:1000].to(torch.int64) * 3).sum().item()      # eager work between replays


def test_topk_large_u32():
    """2^26 keys, k = 10^4: the selection of the serving workload's largest
    requests, against np.partition."""
    n, k = 1 << 26, 10000
    a = datagen.keys_u32(n, 0x5EED0009)
    got = Algorithm.TopK(a, k)
    ref = np.sort(np.partition(a, n - k)[n - k:])[::-1]
    assert np.array_equal(got, ref)
