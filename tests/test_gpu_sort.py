"""GPU parity: libhpcsort.so against the oracle and the golden fixtures.

Bit-exact throughout (integer work).  Host-array calls go through the
reference-shaped surface (hpcsharp_amd.algorithm); device-resident calls go
through the C ABI on torch buffers.  The full-size case (2^28 keys) is
checked by size-independent properties: equality with an independent
library sort of the same keys and byte histograms of input and output.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd import _native as N
from hpcsharp_amd.algorithm import (Algorithm, ArgumentNullException, ArgumentOutOfRangeException,
                                    ParallelAlgorithm)
from tests._fixtures import FIXTURES, float_order, hist_bytes, record_dtype, sha

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    N.context(0)
    yield


def test_reference_examples():
    ex = FIXTURES["reference_examples"]
    a = np.array(ex["program_cs_uint"]["input"], dtype=np.uint32)
    assert Algorithm.SortRadix(a) is a
    assert a.tolist() == ex["program_cs_uint"]["expected"]
    b = np.array(ex["sorting_usage_long"]["input"], dtype=np.uint64)
    assert ParallelAlgorithm.SortRadixPar(b) is b
    assert b.tolist() == ex["sorting_usage_long"]["expected"]
    u = ex["user_defined_class"]
    rec = np.zeros(6, dtype=record_dtype(np.uint32))
    rec["key"] = u["keys"]
    rec["payload"] = u["index"]
    Algorithm.SortRadixLsd2(rec, "key")
    assert rec["key"].tolist() == u["expected_keys"]
    assert rec["payload"].tolist() == u["expected_index"]


@pytest.mark.parametrize("case", FIXTURES["u32"], ids=lambda c: f"{c['dist']}-{c['n']}")
def test_u32_golden_and_oracle(oracle, case):
    a = datagen.keys_u32(case["n"], case["seed"], case["dist"])
    ref = a.copy()
    oracle.sort_radix_par(ref)
    Algorithm.SortRadixLsd(a)
    assert sha(a) == case["sorted_sha"]
    assert np.array_equal(a, ref)


@pytest.mark.parametrize("dist", ["uniform", "low8", "sorted", "constant"])
def test_u32_config1_2pow20(oracle, dist):
    n = 1 << 20
    a = datagen.keys_u32(n, 0x5EED0001, dist)
    ref = a.copy()
    oracle.sort_radix_lsd(ref)
    h_gpu = Algorithm.HistogramByteComponents(a, 0, n - 1)
    h_ref = oracle.histogram_byte_components_u32(a, 0, n - 1)
    assert np.array_equal(h_gpu, h_ref)
    assert np.array_equal(ParallelAlgorithm.HistogramByteComponentsPar(a, 0, n - 1),
                          oracle.histogram_byte_components_par(a, 0, n - 1))
    ParallelAlgorithm.SortRadixPar(a)
    assert np.array_equal(a, ref)


@pytest.mark.parametrize("case", FIXTURES["u64"], ids=lambda c: str(c["n"]))
def test_u64_golden(oracle, case):
    a = datagen.keys_u64(case["n"], case["seed"])
    ref = a.copy()
    oracle.sort_radix_lsd(ref)
    Algorithm.SortRadixLsd(a)
    assert sha(a) == case["sorted_sha"]
    assert np.array_equal(a, ref)


@pytest.mark.parametrize("case", FIXTURES["f64"], ids=lambda c: str(c["n"]))
def test_f64_golden(oracle, case):
    x = datagen.doubles_special(case["n"], case["seed"])
    ref = x.copy()
    oracle.sort_radix_float(ref)
    Algorithm.SortRadixMsd(x)
    assert sha(x) == case["sorted_sha"]
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("case", FIXTURES["f32"], ids=lambda c: str(c["n"]))
def test_f32_golden(oracle, case):
    x = datagen.floats_finite(case["n"], case["seed"])
    ref = x.copy()
    oracle.sort_radix_float(ref)
    Algorithm.SortRadixMsd(x)
    assert sha(x) == case["sorted_sha"]
    assert np.array_equal(x.view(np.uint32), ref.view(np.uint32))


def test_f64_specials():
    sp = FIXTURES["f64_specials"]
    x = np.array([int(h, 16) for h in sp["input_bits"]], dtype=np.uint64).view(np.float64)
    Algorithm.SortRadixMsd(x)
    assert [hex(v) for v in x.view(np.uint64)] == sp["expected_bits"]


def test_f64_all_trivial_but_one_pass(oracle):
    # only the sign differs: every pass but the top one is trivial, so the
    # key transform must enter and leave in that one pass
    x = np.array([1.0, -1.0, 1.0, -1.0, -1.0] * 1000, dtype=np.float64)
    ref = x.copy()
    oracle.sort_radix_float(ref)
    Algorithm.SortRadixMsd(x)
    assert np.array_equal(x.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("case", FIXTURES["pairs"], ids=lambda c: f"{c['n']}-{c['mask']}")
def test_pairs_stable_golden(case):
    k = datagen.keys_u64(case["n"], case["seed"]) & np.uint64(int(case["mask"], 16))
    items = np.arange(case["n"], dtype=np.uint32)
    out_items, out_keys = Algorithm.SortRadix(items, k)
    assert out_items is items and out_keys is k
    assert sha(k) == case["sorted_keys_sha"]
    assert sha(items) == case["perm_sha"]


def test_pairs_u32_stable(oracle):
    n = 300001
    k = datagen.keys_u32(n, 42) & np.uint32(0xFFF)
    v = np.arange(n, dtype=np.uint32)
    rk, rv = k.copy(), v.copy()
    oracle.sort_radix_pairs(rk, rv)
    Algorithm.SortRadix(v, k)
    assert np.array_equal(k, rk) and np.array_equal(v, rv)


@pytest.mark.parametrize("key_type,mask", [(np.uint64, 0xFFFF), (np.uint64, 0xFFFFFFFFFFFFFFFF), (np.uint32, 0xFF)])
def test_records_stable(oracle, key_type, mask):
    n = 200003
    rec = np.zeros(n, dtype=record_dtype(key_type))
    rec["key"] = (datagen.keys_u64(n, 7) & np.uint64(mask)).astype(key_type)
    rec["payload"] = np.arange(n, dtype=np.uint32)
    ref = _byte_copy(rec)
    oracle.sort_radix_records(ref, 0, np.dtype(key_type).itemsize * 8)
    assert Algorithm.SortRadixLsd2(rec, "key") is rec
    _assert_records_equal(rec, ref)


def _byte_copy(rec):
    # ndarray.copy() of a structured dtype with gaps leaves the padding bytes
    # uninitialised; the record sorts move padding too, so copy every byte
    return np.frombuffer(bytearray(rec.tobytes()), dtype=rec.dtype)


def _assert_records_equal(rec, ref):
    if rec.tobytes() == ref.tobytes():
        return
    for f in rec.dtype.names:
        bad = np.flatnonzero(rec[f] != ref[f])
        if bad.size:
            i = int(bad[0])
            raise AssertionError(f"field {f}: {bad.size} records differ, first at {i}: got {rec[i]} want {ref[i]}; "
                                 f"got[{i}-1..{i}+1]={rec[max(0, i - 1):i + 2]} want={ref[max(0, i - 1):i + 2]}")
    rb = np.frombuffer(rec.tobytes(), np.uint8).reshape(rec.size, -1)
    fb = np.frombuffer(ref.tobytes(), np.uint8).reshape(ref.size, -1)
    i = int(np.flatnonzero((rb != fb).any(axis=1))[0])
    raise AssertionError(f"padding bytes differ first at record {i}: got {rb[i].tolist()} want {fb[i].tolist()}")


def test_records_float_key_unaligned(oracle):
    n = 70001
    dt = np.dtype({"names": ["a", "key", "payload"], "formats": [np.uint8, np.float64, np.uint32],
                   "offsets": [0, 1, 9], "itemsize": 13})           # packed: key at an odd offset
    rec = np.zeros(n, dtype=dt)
    rec["key"] = datagen.doubles_special(n, 99)
    rec["payload"] = np.arange(n, dtype=np.uint32)
    ref = _byte_copy(rec)
    oracle.sort_radix_records(ref, 1, 64, 1)
    Algorithm.SortRadixLsd2(rec, "key")
    _assert_records_equal(rec, ref)


def test_records_32_byte_and_work_buffer_reuse(oracle):
    """32-byte records (copied in two 16-byte vectors per record by the key
    extraction) and 12-byte ones (word copy), through one context whose work
    buffer grows and is then reused by a smaller call: the oracle's order,
    padding bytes included, on host arrays and on a device sub-array."""
    import torch
    P = ctypes.c_void_p
    dt32 = np.dtype({"names": ["payload", "key", "tail"], "formats": [np.uint32, np.uint64, "u8,u8"],
                     "offsets": [0, 8, 16], "itemsize": 32})
    dt12 = np.dtype({"names": ["key", "payload"], "formats": [np.uint64, np.uint32], "offsets": [0, 8],
                     "itemsize": 12})
    ctx = N.new_context(0)
    try:
        for n, dt, koff in ((1000, dt32, 8), (300001, dt32, 8), (250007, dt12, 0), (5000, dt32, 8)):
            print(f"records {dt.itemsize} B, n={n}", flush=True)
            rec = np.zeros(n, dtype=dt)
            rec["key"] = datagen.keys_u64(n, 40 + n) & np.uint64(0xFFFF0000000000FF)   # ties: stability
            rec["payload"] = np.arange(n, dtype=np.uint32)
            if "tail" in dt.names:
                rec["tail"]["f0"] = datagen.keys_u64(n, 41 + n)
                rec["tail"]["f1"] = np.arange(n, dtype=np.uint64) * np.uint64(3)
            ref = _byte_copy(rec)
            oracle.sort_radix_records(ref, koff, 64)
            N.check(N.lib().hpcs_sort_radix_records(ctx, rec.ctypes.data_as(P), n, dt.itemsize, koff, 64, 0))
            _assert_records_equal(rec, ref)
        # device call on records that start 32 bytes into a buffer (16-byte aligned)
        n = 70001
        rec = np.zeros(n + 1, dtype=dt32)
        rec["key"] = datagen.keys_u64(n + 1, 77)
        rec["payload"] = np.arange(n + 1, dtype=np.uint32)
        ref = _byte_copy(rec)
        oracle.sort_radix_records(ref[1:], 8, 64)
        t = torch.from_numpy(np.frombuffer(rec.tobytes(), np.uint8).copy()).cuda()
        st = P(torch.cuda.current_stream().cuda_stream)
        N.check(N.lib().hpcs_dev_sort_radix_records(ctx, P(t.data_ptr() + 32), n, 32, 8, 64, 0, st))
        torch.cuda.synchronize()
        got = np.frombuffer(t.cpu().numpy().tobytes(), dtype=dt32)
        _assert_records_equal(got, ref)
    finally:
        N.destroy_context(ctx)


@pytest.mark.parametrize("case", FIXTURES["hist"], ids=lambda c: str(c["n"]))
def test_histograms(case):
    n = case["n"]
    a = datagen.keys_u32(n, case["seed_u32"], "uniform")
    assert sha(Algorithm.HistogramByteComponents(a, 0, n - 1)) == case["u32_sha"]
    b = datagen.keys_u64(n, case["seed_u64"])
    assert sha(Algorithm.HistogramByteComponents(b, 0, n - 1)) == case["u64_sha"]


def test_histogram_subrange_and_types(oracle):
    a = datagen.keys_u32(100000, 3)
    h = Algorithm.HistogramByteComponents(a, 17, 99000)
    assert h.dtype == np.int32
    assert np.array_equal(h, oracle.histogram_byte_components_u32(a, 17, 99000))
    hp = ParallelAlgorithm.HistogramByteComponentsPar(a, 17, 99000)
    assert hp.dtype == np.uint32 and np.array_equal(hp, h.astype(np.uint32))


def test_sub_range_sort(oracle):
    a = datagen.keys_u32(200000, 11)
    ref = a.copy()
    oracle.sort_radix_lsd(ref, 12345, 150000)
    Algorithm.SortRadixLsd(a, 12345, 150000)
    assert np.array_equal(a, ref)


def test_argument_errors():
    with pytest.raises(ArgumentNullException):
        Algorithm.SortRadixLsd(None)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.SortRadixLsd(np.zeros(10, dtype=np.uint32), 5, 10)
    rc = N.lib().hpcs_sort_radix_records(N.context(), None, 10, 16, 12, 64, 0)
    assert rc == N.HPCS_E_INVALID_ARG


def _torch_u32(a: np.ndarray):
    import torch
    return torch.from_numpy(a.view(np.int32)).cuda()


def test_device_api_u32_and_stream(oracle):
    import torch
    n = 3 * (1 << 20) + 12345
    a = datagen.keys_u32(n, 1234)
    ref = a.copy()
    oracle.sort_radix_lsd(ref)
    t = _torch_u32(a)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        N.check(N.lib().hpcs_dev_sort_radix_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                                ctypes.c_void_p(s.cuda_stream)))
    s.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ref)
    counts = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                                           ctypes.c_void_p(counts.data_ptr()), None))
    torch.cuda.synchronize()
    N.check(N.lib().hpcs_ctx_trim(N.context()))
    assert np.array_equal(counts.cpu().numpy().reshape(4, 256), hist_bytes(a))


def test_full_size_2pow28_properties():
    """2^28 uniform u32 (the headline size): the output equals an independent
    library sort of the same keys (torch.sort on an order-preserving int64
    copy), and input and output have the same byte histograms."""
    import torch
    n = 1 << 28
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)   # order with torch's own work
    g = torch.Generator(device="cuda")
    g.manual_seed(0x5EED0002)
    t = torch.randint(-(1 << 31), 1 << 31, (n,), dtype=torch.int32, device="cuda", generator=g)
    h_in = torch.empty(1024, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                                           ctypes.c_void_p(h_in.data_ptr()), st))
    ref = torch.sort((t.to(torch.int64) & 0xFFFFFFFF))[0]
    N.check(N.lib().hpcs_dev_sort_radix_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n, st))
    got = t.to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(got, ref)
    del ref, got
    h_out = torch.empty(1024, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                                           ctypes.c_void_p(h_out.data_ptr()), st))
    torch.cuda.synchronize()
    assert torch.equal(h_in, h_out)
    assert int(h_in.view(4, 256).sum(dim=1).min()) == n


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 6, 7, 8, 9])
def test_forced_lookback_fallback(oracle, shape):
    """Every look-back wait resolved by counting the earlier tile locally
    (fallback threshold 0) still gives the oracle's output, on every tile
    shape (a shape that does not serve a key type leaves it on its default)."""
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, 0, ctx)          # single-pass look-back
        N.set_option(N.HPCS_OPT_FALLBACK_ROUNDS, 0, ctx)
        N.set_option(N.HPCS_OPT_TILE_SHAPE, shape, ctx)
        a = datagen.keys_u32((1 << 20) + 333, 4242 + shape)
        ref = a.copy()
        oracle.sort_radix_lsd(ref)
        N.check(N.lib().hpcs_sort_radix_u32(ctx, a.ctypes.data_as(ctypes.c_void_p), a.size))
        assert np.array_equal(a, ref)
        k = datagen.keys_u64(300001, 77) & np.uint64(0xFFFFF)
        v = np.arange(k.size, dtype=np.uint32)
        rk, rv = k.copy(), v.copy()
        oracle.sort_radix_pairs(rk, rv)
        N.check(N.lib().hpcs_sort_radix_pairs_u64_u32(ctx, k.ctypes.data_as(ctypes.c_void_p),
                                                      v.ctypes.data_as(ctypes.c_void_p), k.size))
        assert np.array_equal(k, rk) and np.array_equal(v, rv)
        x = datagen.doubles_special(200003, 5)
        xr = x.copy()
        oracle.sort_radix_float(xr)
        N.check(N.lib().hpcs_sort_radix_f64(ctx, x.ctypes.data_as(ctypes.c_void_p), x.size))
        assert np.array_equal(x.view(np.uint64), xr.view(np.uint64))
    finally:
        N.destroy_context(ctx)


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 7, 8, 9])
def test_tile_shapes_u32(oracle, shape, algo):
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_TILE_SHAPE, shape, ctx)
        N.set_option(N.HPCS_OPT_PASS_ALGO, algo, ctx)
        for n, dist in [(8191, "uniform"), (8192, "sorted"), (8193, "low8"), (3 * (1 << 20) + 7, "uniform")]:
            a = datagen.keys_u32(n, 99 + n, dist)
            ref = a.copy()
            oracle.sort_radix_lsd(ref)
            N.check(N.lib().hpcs_sort_radix_u32(ctx, a.ctypes.data_as(ctypes.c_void_p), a.size))
            assert np.array_equal(a, ref), (n, dist)
    finally:
        N.destroy_context(ctx)


@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("shape", [4, 6])
def test_small_tile_shape_eight_byte_keys(oracle, algo, shape):
    """The 256x8 tile shape (any sort) and the 1024x8 shape (8-byte keys,
    4-byte keys with values) give the oracle's output in both pass
    algorithms, for 8-byte keys and for pairs."""
    ctx = N.new_context(0)
    try:
        N.set_option(N.HPCS_OPT_TILE_SHAPE, shape, ctx)
        N.set_option(N.HPCS_OPT_PASS_ALGO, algo, ctx)
        for n in (2047, 2048 * 3 + 1, (1 << 21) + 5):
            a = datagen.keys_u64(n, 7 + n)
            ref = a.copy()
            oracle.sort_radix_lsd(ref)
            N.check(N.lib().hpcs_sort_radix_u64(ctx, a.ctypes.data_as(ctypes.c_void_p), a.size))
            assert np.array_equal(a, ref), n
            k = datagen.keys_u64(n, 11 + n) & np.uint64(0xFF00FF)
            v = np.arange(n, dtype=np.uint32)
            order = np.argsort(k, kind="stable")
            N.check(N.lib().hpcs_sort_radix_pairs_u64_u32(ctx, k.ctypes.data_as(ctypes.c_void_p),
                                                          v.ctypes.data_as(ctypes.c_void_p), n))
            assert np.array_equal(v, order.astype(np.uint32)), n
            k32 = datagen.keys_u32(n, 13 + n) & np.uint32(0xF0F0F)
            v32 = np.arange(n, dtype=np.uint32)
            order32 = np.argsort(k32, kind="stable")
            N.check(N.lib().hpcs_sort_radix_pairs_u32_u32(ctx, k32.ctypes.data_as(ctypes.c_void_p),
                                                          v32.ctypes.data_as(ctypes.c_void_p), n))
            assert np.array_equal(v32, order32.astype(np.uint32)), n
    finally:
        N.destroy_context(ctx)


@pytest.mark.parametrize("mode", ["bytes", "keys"])
def test_rts_next_digit_bytes(oracle, mode):
    """Reduce-then-scan passes count their tiles from the digit bytes the
    previous moving pass wrote (or from the keys): the same output as the
    oracle when the first pass, or passes in the middle, move nothing."""
    ctx = N.new_context(0)
    P = ctypes.c_void_p
    try:
        N.set_option(N.HPCS_OPT_PASS_ALGO, 1, ctx)
        with pytest.raises(N.HpcsError):
            N.set_option(N.HPCS_OPT_DIGIT_BYTES, 3, ctx)
        N.set_option(N.HPCS_OPT_DIGIT_BYTES, 0 if mode == "keys" else 1, ctx)
        n = 3 * (1 << 20) + 4097                    # a ragged last tile
        base = datagen.keys_u32(n, 808)
        for mask in (0xFFFFFFFF, 0xFFFFFF00, 0x00FF00FF, 0xFF000000, 0x0000FF00):
            a = base & np.uint32(mask)
            ra = a.copy()
            oracle.sort_radix_lsd(ra)
            N.check(N.lib().hpcs_sort_radix_u32(ctx, a.ctypes.data_as(P), n))
            assert np.array_equal(a, ra), hex(mask)
        for mask in (0xFFFFFFFFFFFFFFFF, 0xFF00FF00FF00FF00, 0x00000000FFFF0000):
            b = datagen.keys_u64(n, 809) & np.uint64(mask)
            rb = b.copy()
            oracle.sort_radix_lsd(rb)
            N.check(N.lib().hpcs_sort_radix_u64(ctx, b.ctypes.data_as(P), n))
            assert np.array_equal(b, rb), hex(mask)
        x = datagen.doubles_special(n, 810)
        rx = x.copy()
        oracle.sort_radix_float(rx)
        N.check(N.lib().hpcs_sort_radix_f64(ctx, x.ctypes.data_as(P), n))
        assert np.array_equal(x.view(np.uint64), rx.view(np.uint64))
        k = datagen.keys_u64(n, 811) & np.uint64(0xFFFF00FF00)
        v = np.arange(n, dtype=np.uint32)
        order = np.argsort(k, kind="stable").astype(np.uint32)
        N.check(N.lib().hpcs_sort_radix_pairs_u64_u32(ctx, k.ctypes.data_as(P), v.ctypes.data_as(P), n))
        assert np.array_equal(v, order)
    finally:
        N.destroy_context(ctx)


def test_both_pass_algorithms_all_types(oracle):
    """The single-pass look-back form and the reduce-then-scan form give the
    oracle's output for every key type and for key-value pairs."""
    for algo in (0, 1):
        ctx = N.new_context(0)
        try:
            N.set_option(N.HPCS_OPT_PASS_ALGO, algo, ctx)
            P = ctypes.c_void_p
            a = datagen.keys_u64(400009, 31 + algo)
            ra = a.copy()
            oracle.sort_radix_lsd(ra)
            N.check(N.lib().hpcs_sort_radix_u64(ctx, a.ctypes.data_as(P), a.size))
            assert np.array_equal(a, ra)
            x = datagen.floats_finite(300007, 41 + algo)
            rx = x.copy()
            oracle.sort_radix_float(rx)
            N.check(N.lib().hpcs_sort_radix_f32(ctx, x.ctypes.data_as(P), x.size))
            assert np.array_equal(x.view(np.uint32), rx.view(np.uint32))
            k = datagen.keys_u32(250001, 51 + algo) & np.uint32(0x3FFF)
            v = np.arange(k.size, dtype=np.uint32)
            rk, rv = k.copy(), v.copy()
            oracle.sort_radix_pairs(rk, rv)
            N.check(N.lib().hpcs_sort_radix_pairs_u32_u32(ctx, k.ctypes.data_as(P), v.ctypes.data_as(P), k.size))
            assert np.array_equal(k, rk) and np.array_equal(v, rv)
            for dist in ("sorted", "low8", "constant"):
                b = datagen.keys_u32(1 << 20, 61, dist)
                rb = b.copy()
                oracle.sort_radix_lsd(rb)
                N.check(N.lib().hpcs_sort_radix_u32(ctx, b.ctypes.data_as(P), b.size))
                assert np.array_equal(b, rb), dist
        finally:
            N.destroy_context(ctx)


def _segments(nseg, seed, lo, hi, gaps=False):
    """Log-uniform lengths with some empty and single-key segments; offsets
    are the running sum (segments tile [0, n))."""
    lens = datagen.segment_lengths(nseg, seed, lo, hi)
    rng = np.random.default_rng(seed)
    lens[rng.integers(0, nseg, size=max(1, nseg // 20))] = 0          # empty segments
    lens[rng.integers(0, nseg, size=max(1, nseg // 20))] = 1          # single keys
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return begin, lens


@pytest.mark.parametrize("nseg,lo,hi", [(300, 16, 1 << 13), (40, 4000, 1 << 20), (7, 1, 9000)])
def test_segments_host(oracle, nseg, lo, hi):
    begin, lens = _segments(nseg, 1000 + nseg, lo, hi, gaps=False)
    n = int(begin[-1])
    a = datagen.keys_u32(n + 5, 77 + nseg)
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    N.check(N.lib().hpcs_sort_segments_u32(N.context(), a.ctypes.data_as(ctypes.c_void_p),
                                           begin.ctypes.data_as(ctypes.c_void_p), nseg))
    assert np.array_equal(a, ref)


def test_segments_device_with_gaps_and_distributions(oracle):
    import torch
    _, lens = _segments(500, 4242, 2, 1 << 15, gaps=False)
    # three untouched keys on either side of the batch
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64) + 3
    n = int(begin[-1]) + 3
    a = datagen.keys_u32(n, 99)
    a[: n // 3] &= np.uint32(0xFF)               # low-8 keys in the first third
    a[n // 3: n // 2] = np.arange(n // 2 - n // 3, dtype=np.uint32)   # sorted run
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    t = torch.from_numpy(a.view(np.int32)).cuda()
    b = torch.from_numpy(begin).cuda()
    st = torch.cuda.current_stream().cuda_stream
    N.check(N.lib().hpcs_dev_sort_segments_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                               ctypes.c_void_p(b.data_ptr()), len(lens), ctypes.c_void_p(st)))
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ref)


@pytest.mark.parametrize("m", [150, 200000])
def test_segments_device_rejects_overlapping_offsets(m):
    """Offsets that are not non-decreasing, [0, m, 0, m, ...] and [0, m, m/2,
    3m/2, 2m, 2m + 500]: only segments within [0, n) that start at or after
    the end of every earlier valid segment are sorted; every other key,
    and a canary past the keys, is untouched, and the same context sorts a
    normal batch correctly afterwards."""
    import torch
    n = 2 * m + 1000
    canary = 64
    a = datagen.keys_u32(n + canary, 0x5EED0400 + m)
    for begin in (np.array([0, m, 0, m, 0, m, 0, m], dtype=np.int64),
                  np.array([0, m, m // 2, m + m // 2, 2 * m, 2 * m + 500], dtype=np.int64)):
        nseg = begin.size - 1
        want = a.copy()
        want[:m] = np.sort(a[:m])
        if begin.size == 6:
            # [m/2, 3m/2) overlaps [0, m) and is skipped; [3m/2, 2m) and
            # [2m, 2m + 500) start at the end of an earlier segment: sorted
            want[m + m // 2:2 * m] = np.sort(a[m + m // 2:2 * m])
            want[2 * m:2 * m + 500] = np.sort(a[2 * m:2 * m + 500])
        t = torch.from_numpy(a.view(np.int32).copy()).cuda()
        b = torch.from_numpy(begin).cuda()
        st = torch.cuda.current_stream().cuda_stream
        N.check(N.lib().hpcs_dev_sort_segments_u32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                                   ctypes.c_void_p(b.data_ptr()), nseg, ctypes.c_void_p(st)))
        torch.cuda.synchronize()
        got = t.cpu().numpy().view(np.uint32)
        assert np.array_equal(got[:n], want[:n]), begin.tolist()
        assert np.array_equal(got[n:], a[n:]), "canary past the keys"
    lens = datagen.segment_lengths(64, 3, 16, 1 << 16)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    k = datagen.keys_u32(int(begin[-1]), 4)
    want = k.copy()
    for s0, s1 in zip(begin[:-1], begin[1:]):
        want[s0:s1] = np.sort(want[s0:s1])
    t = torch.from_numpy(k.view(np.int32).copy()).cuda()
    b = torch.from_numpy(begin).cuda()
    N.check(N.lib().hpcs_dev_sort_segments_u32(N.context(), ctypes.c_void_p(t.data_ptr()), int(begin[-1]),
                                               ctypes.c_void_p(b.data_ptr()), len(lens),
                                               ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint32), want)


def test_segments_graph_replays_match_eager(oracle):
    """The batched call captured once in a HIP graph: every replay (with eager
    work between replays) equals the eager result."""
    import torch
    lens = datagen.segment_lengths(256, 5, 16, 1 << 17)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(begin[-1])
    a = datagen.keys_u32(n, 8)
    ref = a.copy()
    oracle.sort_segments_u32(ref, begin)
    src = torch.from_numpy(a.view(np.int32)).cuda()
    work = src.clone()
    b = torch.from_numpy(begin).cuda()
    ctx = N.context()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        N.check(N.lib().hpcs_dev_sort_segments_u32(ctx, ctypes.c_void_p(work.data_ptr()), n,
                                                   ctypes.c_void_p(b.data_ptr()), len(lens),
                                                   ctypes.c_void_p(s.cuda_stream)))   # warm-up sizes scratch
    s.synchronize()
    assert np.array_equal(work.cpu().numpy().view(np.uint32), ref)
    g = torch.cuda.CUDAGraph()
    work.copy_(src)
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        N.check(N.lib().hpcs_dev_sort_segments_u32(ctx, ctypes.c_void_p(work.data_ptr()), n,
                                                   ctypes.c_void_p(b.data_ptr()), len(lens),
                                                   ctypes.c_void_p(s.cuda_stream)))
    for rep in range(3):
        work.copy_(src)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(work.cpu().numpy().view(np.uint32), ref), rep
        _ = (src[:1000].to(torch.int64) * 3).sum().item()    # eager work between replays


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_device_unaligned_subarray(oracle, offset):
    """Device calls on a pointer that is not 16-byte aligned (a sub-array)."""
    import torch
    n = (1 << 20) + 77
    a = datagen.keys_u32(n + 8, 300 + offset)
    ref = a.copy()
    oracle.sort_radix_lsd(ref, offset, n)
    t = torch.from_numpy(a.view(np.int32)).cuda()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    N.check(N.lib().hpcs_dev_sort_radix_u32(N.context(), ctypes.c_void_p(t.data_ptr() + 4 * offset), n, st))
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy().view(np.uint32), ref)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("n", [1000, 70001, (1 << 20) + 3])
def test_signed_keys_vs_oracle(oracle, dtype, n):
    """int[] / long[]: sort, parallel sort, histogram (top byte ^ 128) and
    top-k against the oracle's restatements (small n exercises the onesweep
    pass, the largest the reduce-then-scan pass)."""
    x = datagen.keys_u64(n, 900 + n).view(np.int64)
    x = (x >> np.int64(32)).astype(np.int32) if dtype == np.int32 else x.copy()
    x[:4] = [0, -1, np.iinfo(dtype).min, np.iinfo(dtype).max]
    ref = x.copy()
    oracle.sort_radix_signed(ref)
    h = Algorithm.HistogramByteComponents(x, 0, n - 1)
    assert np.array_equal(h, oracle.histogram_byte_components_signed(x, 0, n - 1))
    top = Algorithm.TopK(x, 777)
    assert np.array_equal(top, ref[::-1][:777])
    a = x.copy()
    Algorithm.SortRadixLsd(a)
    assert np.array_equal(a, ref)
    b = x.copy()
    assert ParallelAlgorithm.SortRadixPar(b) is b
    assert np.array_equal(b, ref)


def test_signed_device_and_records(oracle):
    import torch
    n = 300007
    x = (datagen.keys_u64(n, 77).view(np.int64) >> np.int64(40)).astype(np.int32)
    ref = x.copy()
    oracle.sort_radix_signed(ref)
    t = torch.from_numpy(x.copy()).cuda()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    N.check(N.lib().hpcs_dev_sort_radix_i32(N.context(), ctypes.c_void_p(t.data_ptr()), n, st))
    counts = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_i32(N.context(), ctypes.c_void_p(t.data_ptr()), n,
                                                           ctypes.c_void_p(counts.data_ptr()), st))
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), ref)
    assert np.array_equal(counts.cpu().numpy().reshape(4, 256), oracle.histogram_byte_components_signed(x, 0, n - 1))
    rec = np.zeros(n, dtype=record_dtype(np.uint64))
    keys = (datagen.keys_u64(n, 78) & np.uint64(0xFFFFF)).astype(np.int64) - (1 << 19)
    rec["key"] = keys.view(np.uint64)
    rec["payload"] = np.arange(n, dtype=np.uint32)
    want = rec.copy()
    oracle.sort_radix_records(want, 0, 64, oracle.KEY_INT)
    rc = N.lib().hpcs_sort_radix_records(N.context(), rec.ctypes.data_as(ctypes.c_void_p), n, rec.dtype.itemsize,
                                         0, 64, N.HPCS_KEY_INT)
    N.check(rc)
    assert np.array_equal(rec["key"], want["key"]) and np.array_equal(rec["payload"], want["payload"])


@pytest.mark.parametrize("key_type", [np.uint32, np.uint64])
def test_key_extractor_histogram(oracle, key_type):
    """HistogramByteComponents<T>(T[], l, r, getKey): the byte counts of the
    records' keys over [l, r] (brute force), host and device forms."""
    import torch
    n = 100003
    rec = np.zeros(n, dtype=np.dtype({"names": ["pad", "key", "payload"], "formats": [np.uint8, key_type, np.uint32],
                                      "offsets": [0, 4 if key_type == np.uint32 else 8, 16], "itemsize": 24}))
    rec["key"] = datagen.keys_u64(n, 515).astype(key_type)
    l, r = 17, n - 30
    h = Algorithm.HistogramByteComponents(rec, l, r, "key")
    assert np.array_equal(h, hist_bytes(np.ascontiguousarray(rec["key"][l:r + 1])))
    d = torch.from_numpy(rec.view(np.uint8).reshape(-1).copy()).cuda()
    counts = torch.empty(np.dtype(key_type).itemsize * 256, dtype=torch.int32, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    N.check(N.lib().hpcs_dev_histogram_byte_components_records(
        N.context(), ctypes.c_void_p(d.data_ptr()), n, 24, rec.dtype.fields["key"][1], 8 * np.dtype(key_type).itemsize,
        ctypes.c_void_p(counts.data_ptr()), st))
    torch.cuda.synchronize()
    assert np.array_equal(counts.cpu().numpy().reshape(-1, 256), hist_bytes(np.ascontiguousarray(rec["key"])))


def test_tuning_table_selects_the_variant(oracle, tmp_path):
    """A tuning row picks the tile shape, pass algorithm and ranking for its
    size class; an option set explicitly wins; the output never changes."""
    ctx = N.new_context(0)
    try:
        p = tmp_path / "tune.txt"
        p.write_text("u32 2097152 0 3 0\nu32 268435456 1 0 1\n")
        N.load_tuning(str(p), ctx)
        info = N.PlanInfo()
        N.check(N.lib().hpcs_plan_query(ctx, N.HPCS_DTYPE_U32, 1 << 20, 0, ctypes.byref(info)))
        assert info.tile_keys == 512 * 32
        N.check(N.lib().hpcs_plan_query(ctx, N.HPCS_DTYPE_U32, 1 << 24, 0, ctypes.byref(info)))
        assert info.tile_keys == 256 * 16
        n = (1 << 20) + 999
        a = datagen.keys_u32(n, 4321)
        ref = a.copy()
        oracle.sort_radix_lsd(ref)
        N.check(N.lib().hpcs_sort_radix_u32(ctx, a.ctypes.data_as(ctypes.c_void_p), n))
        assert np.array_equal(a, ref)
        N.set_option(N.HPCS_OPT_TILE_SHAPE, 1, ctx)       # explicit option beats the table
        N.check(N.lib().hpcs_plan_query(ctx, N.HPCS_DTYPE_U32, 1 << 20, 0, ctypes.byref(info)))
        assert info.tile_keys == 512 * 16
        N.load_tuning(None, ctx)
        b = datagen.keys_u32(n, 4322)
        ref = b.copy()
        oracle.sort_radix_lsd(ref)
        N.check(N.lib().hpcs_sort_radix_u32(ctx, b.ctypes.data_as(ctypes.c_void_p), n))
        assert np.array_equal(b, ref)
    finally:
        N.destroy_context(ctx)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
@pytest.mark.parametrize("bits", [1, 5, 8, 11, 12, 16])
def test_nbit_histograms_vs_oracle(oracle, dtype, bits):
    """Histogram11bitComponents / HistogramNbitComponents against the oracle's
    restatement, over a sub-range (LDS tables up to 64 KiB, global beyond)."""
    n = 200003
    a = datagen.keys_u64(n, 70 + bits).astype(dtype)
    a[:1000] = a[0]                                    # a run of equal keys (one add per wave)
    l, r = 3, n - 11
    got = Algorithm.HistogramNbitComponents(a, l, r, bits)
    assert np.array_equal(got, oracle.histogram_nbit_components(a, l, r, bits))
    if dtype == np.uint32 and bits == 11:
        assert np.array_equal(Algorithm.Histogram11bitComponents(a, l, r), got)


@pytest.mark.parametrize("key_type", [np.uint32, np.uint64])
def test_items_by_separate_keys(oracle, key_type):
    """SortRadix<T>(T[] items, keys[]): items of any size move with their
    keys, stably (narrow keys give many ties; payload = original index)."""
    n = 150001
    keys = (datagen.keys_u64(n, 31) & np.uint64(0x3FF)).astype(key_type)
    items = np.zeros(n, dtype=np.dtype({"names": ["idx", "tag"], "formats": [np.uint32, "S9"], "itemsize": 13}))
    items["idx"] = np.arange(n, dtype=np.uint32)
    items["tag"] = [b"x%08d" % i for i in range(n)]
    want = np.argsort(keys, kind="stable")
    k2, v2 = keys.copy(), np.arange(n, dtype=np.uint32)
    oracle.sort_radix_pairs(k2.astype(np.uint64) if key_type == np.uint64 else k2, v2)
    assert np.array_equal(v2, want.astype(np.uint32))              # the oracle agrees with the closed form
    got_items, got_keys = Algorithm.SortRadix(items, keys)
    assert np.array_equal(got_keys, np.sort(keys.copy()))
    assert np.array_equal(got_items["idx"], want.astype(np.uint32))
    assert np.array_equal(got_items["tag"], np.array([b"x%08d" % i for i in want]))


def test_beyond_2pow30_keys():
    """n = 2^30 + 12345 (past the look-back words' 30-bit counts): sortedness
    and the multiset (byte histograms, key sum) of a device-generated array."""
    import torch
    n = (1 << 30) + 12345
    g = torch.Generator(device="cuda").manual_seed(30)
    t = torch.randint(-2 ** 31, 2 ** 31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    ctx = N.context()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    h0 = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_u32(ctx, ctypes.c_void_p(t.data_ptr()), n,
                                                           ctypes.c_void_p(h0.data_ptr()), st))
    s0 = int((t.to(torch.int64) & 0xFFFFFFFF).sum().item())
    N.check(N.lib().hpcs_dev_sort_radix_u32(ctx, ctypes.c_void_p(t.data_ptr()), n, st))
    h1 = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_u32(ctx, ctypes.c_void_p(t.data_ptr()), n,
                                                           ctypes.c_void_p(h1.data_ptr()), st))
    torch.cuda.synchronize()
    assert torch.equal(h0, h1)
    u = t.to(torch.int64) & 0xFFFFFFFF
    assert int(u.sum().item()) == s0
    assert bool((u[1:] >= u[:-1]).all().item())
    del u, t
    torch.cuda.empty_cache()
    N.check(N.lib().hpcs_ctx_trim(ctx))


def test_partition_radix_and_msd_entry_points():
    """PartitionRadix (PartitionRadix.cs:135-210): a[k] is the k-th smallest
    of the range, nothing before it larger, nothing after it smaller, the
    range a permutation, the rest untouched; SortRadixMsd for the integer
    key types (RadixSortMsd.cs:682-838, 1285, 1308) sorts."""
    n = 300007
    a = datagen.keys_u32(n, 4040)
    orig = a.copy()
    k = 123456
    Algorithm.PartitionRadix(a, k)
    assert a[k] == np.partition(orig, k)[k]
    assert a[:k].max() <= a[k] <= a[k + 1:].min()
    assert np.array_equal(np.sort(a), np.sort(orig))
    b = datagen.keys_u32(n, 4041)
    ob = b.copy()
    ks = [10, 5000, 77777]
    Algorithm.PartitionRadix(b, 1000, 200000, ks)
    r = b[1000:201000]
    ref = np.sort(ob[1000:201000])
    for kk in ks:
        assert r[kk] == ref[kk]
        assert r[:kk].max(initial=0) <= r[kk] <= r[kk + 1:].min(initial=0xFFFFFFFF)
    assert np.array_equal(b[:1000], ob[:1000]) and np.array_equal(b[201000:], ob[201000:])
    assert np.array_equal(np.sort(r), ref)
    for dt, gen in ((np.uint32, datagen.keys_u32), (np.uint64, datagen.keys_u64),
                    (np.int32, datagen.keys_u32), (np.int64, datagen.keys_u64)):
        x = gen(70001, 4050).view(dt).copy()
        want = np.sort(x)
        Algorithm.SortRadixMsd(x)
        assert np.array_equal(x, want), dt
    y = datagen.keys_u32(50000, 4060)
    oy = y.copy()
    Algorithm.SortRadixMsd(y, 100, 40000, 64)
    assert np.array_equal(y[100:40100], np.sort(oy[100:40100])) and np.array_equal(y[:100], oy[:100])
