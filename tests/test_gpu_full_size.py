"""GPU parity at the benchmark workloads' full sizes, through properties that
do not depend on size and against independent sorts:

- 2^28 u32 in the low-8-bit, ascending and all-equal distributions: equal to
  torch.sort of an order-preserving int64 copy, byte histograms unchanged;
- 2^27 doubles with 1% +-0.0, 1% subnormals and 1% +-1e300: bit-exact
  against the oracle's float sort on the host;
- 2^27 records {u64 key, u32 payload = index} by the 64-bit key: keys in
  order, payloads a permutation that carries each record's key, equal keys
  in payload order (stability), also with 16-bit keys (many ties);
- 4096 u32 segments with lengths log-uniform in [16, 2^20] (3.78e8 keys):
  each segment equal to torch.sort of (segment, key) pairs;
- serving requests of 2^25-2^26 keys: checksums equal the oracle's results,
  alone and run concurrently.

Inputs are generated on the device (hpcsharp_amd.datagen's torch form of the
same splitmix64 streams) so that each case takes seconds.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from hpcsharp_amd import _native as N
from hpcsharp_amd import datagen
from hpcsharp_amd import serving as S

pytestmark = pytest.mark.gpu
P = ctypes.c_void_p
SIGN64 = -(1 << 63)


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    yield
    torch.cuda.empty_cache()


def _stream():
    import torch
    return P(torch.cuda.current_stream().cuda_stream)


def _hist_u32(t):
    import torch
    h = torch.empty(1024, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_histogram_byte_components_u32(N.context(), P(t.data_ptr()), t.numel(),
                                                           P(h.data_ptr()), _stream()))
    return h


@pytest.mark.parametrize("dist", ["low8", "sorted", "constant"])
def test_u32_2pow28_distributions(dist):
    import torch
    n = 1 << 28
    t = datagen.keys_u32_torch(n, 0x5EED0003, dist)
    h_in = _hist_u32(t)
    ref = torch.sort(t.to(torch.int64) & 0xFFFFFFFF)[0]
    N.check(N.lib().hpcs_dev_sort_radix_u32(N.context(), P(t.data_ptr()), n, _stream()))
    assert torch.equal(t.to(torch.int64) & 0xFFFFFFFF, ref), dist
    del ref
    assert torch.equal(_hist_u32(t), h_in)
    torch.cuda.synchronize()


def _doubles_special_torch(n: int, seed: int):
    """N(0, 1e3) doubles (torch's normal generator) with 1% +-0.0, 1%
    subnormals and 1% +-1e300 at disjoint positions chosen, with their signs
    and mantissas, from the splitmix64 stream datagen.doubles_special uses;
    on the device (float64 bits in an int64 tensor)."""
    import torch
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn(n, dtype=torch.float64, device="cuda", generator=g) * 1e3
    bits = x.view(torch.int64).clone()
    r = datagen.splitmix64_torch(seed ^ 0xA5A5A5A5, n)
    sel = (torch.remainder(r, 100) + (r < 0).to(torch.int64) * 16) % 100   # the uint64 value mod 100 (2^64 = 16 mod 100)
    neg = ((r >> 8) & 1) == 1
    zero, sub, big = sel == 0, sel == 1, sel == 2
    bits[zero] = torch.where(neg[zero], SIGN64, 0)
    mant = (r >> 9) & ((1 << 52) - 1)
    mant[mant == 0] = 1
    bits[sub] = torch.where(neg[sub], mant[sub] | SIGN64, mant[sub])
    pos_big = int(np.array([1e300]).view(np.int64)[0])
    bits[big] = torch.where(neg[big], pos_big | SIGN64, pos_big)
    return bits


def test_f64_2pow27_special_values_bit_exact(oracle):
    import torch
    n = 1 << 27
    bits = _doubles_special_torch(n, 0x5EED0006)
    host = bits.cpu().numpy().view(np.float64)
    assert (host == 0).sum() > n // 200 and np.signbit(host[host == 0]).any()
    want = host.copy()
    oracle.sort_radix_float(want)
    N.check(N.lib().hpcs_dev_sort_radix_f64(N.context(), P(bits.data_ptr()), n, _stream()))
    got = bits.cpu().numpy()
    assert np.array_equal(got, want.view(np.int64))
    # -0.0 (radix key 0x7FF...F) before +0.0 (0x800...0), subnormals between
    z = np.nonzero(want == 0)[0]
    assert np.all(np.diff(z) == 1) and np.signbit(want[z[0]]) and not np.signbit(want[z[-1]])


def _u64_order_key(t):
    """int64 values that order as the uint64 bits do."""
    return t ^ SIGN64


@pytest.mark.parametrize("key_mask", [-1, 0xFFFF])
def test_records_2pow27_stable_permutation(key_mask):
    import torch
    n = 1 << 27
    keys = datagen.splitmix64_torch(0x5EED0007, n) & key_mask
    rec = torch.zeros((n, 4), dtype=torch.int32, device="cuda")      # {u64 key, u32 payload, pad}
    rec[:, 0] = (keys & 0xFFFFFFFF).to(torch.int32)
    rec[:, 1] = (keys >> 32).to(torch.int32)
    rec[:, 2] = torch.arange(n, dtype=torch.int32, device="cuda")
    N.check(N.lib().hpcs_dev_sort_radix_records(N.context(), P(rec.data_ptr()), n, 16, 0, 64, N.HPCS_KEY_UINT,
                                                _stream()))
    out_keys = (rec[:, 0].to(torch.int64) & 0xFFFFFFFF) | (rec[:, 1].to(torch.int64) << 32)
    payload = rec[:, 2].to(torch.int64)
    ok = _u64_order_key(out_keys)
    assert bool((ok[1:] >= ok[:-1]).all()), "keys out of order"
    assert torch.equal(torch.sort(payload)[0], torch.arange(n, dtype=torch.int64, device="cuda")), "not a permutation"
    assert torch.equal(keys[payload], out_keys), "a payload moved without its key"
    ties = out_keys[1:] == out_keys[:-1]
    assert bool((payload[1:][ties] > payload[:-1][ties]).all()), "equal keys out of input order"
    if key_mask == 0xFFFF:
        assert int(ties.sum()) > n // 2
    assert bool((rec[:, 3] == 0).all())


def test_segments_4096_log_uniform_full_size():
    import torch
    nseg, seed = 4096, 0x5EED0008
    lens = datagen.segment_lengths(nseg, seed, 16, 1 << 20)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(begin[-1])
    assert total > 3 * 10 ** 8
    t = datagen.keys_u32_torch(total, seed + 1)
    seg = torch.repeat_interleave(torch.arange(nseg, dtype=torch.int64, device="cuda"),
                                  torch.from_numpy(lens).cuda())
    want = torch.sort((seg << 32) | (t.to(torch.int64) & 0xFFFFFFFF))[0] & 0xFFFFFFFF
    del seg
    d_begin = torch.from_numpy(begin).cuda()
    N.check(N.lib().hpcs_dev_sort_segments_u32(N.context(), P(t.data_ptr()), total, P(d_begin.data_ptr()), nseg,
                                               _stream()))
    assert torch.equal(t.to(torch.int64) & 0xFFFFFFFF, want)


def _expected(oracle, req) -> int:
    x = S.payload(req)
    if int(req["op"]) == N.HPCS_OP_SORT:
        (oracle.sort_radix_float if x.dtype.kind == "f" else oracle.sort_radix_lsd)(x)
        return S.checksum(x)
    k = int(req["k"])
    if x.dtype.kind == "f":
        u = x.view(np.uint32)
        t = np.where(u >> np.uint32(31), ~u, u | np.uint32(0x80000000))
        top = x[np.argsort(t, kind="stable")][::-1][:k]
    else:
        top = np.sort(x)[::-1][:k]
    return S.checksum(np.ascontiguousarray(top))


def test_serving_requests_of_2pow25_to_2pow26_keys(oracle):
    """Large serving requests (the workload's top size class) run concurrently
    on four slots and alone: every checksum equals the oracle's result."""
    t = np.zeros(6, dtype=S.REQUEST)
    spec = [(N.HPCS_OP_SORT, N.HPCS_DTYPE_U32, 1 << 26, 0), (N.HPCS_OP_TOPK, N.HPCS_DTYPE_F32, 1 << 25, 10 ** 4),
            (N.HPCS_OP_SORT, N.HPCS_DTYPE_U64, 1 << 25, 0), (N.HPCS_OP_SORT, N.HPCS_DTYPE_F32, (1 << 26) - 3, 0),
            (N.HPCS_OP_TOPK, N.HPCS_DTYPE_U64, 1 << 26, 10), (N.HPCS_OP_TOPK, N.HPCS_DTYPE_U32, (1 << 25) + 1, 7777)]
    for i, (op, dt, n, k) in enumerate(spec):
        t[i] = (op, dt, n, k, 0x5EED0500 + i, 0.0005 * i)
    mix, _ = S.run(t, slots=4)
    alone, _ = S.run(t, mode=N.HPCS_SERVE_ALONE)
    assert np.all(mix["status"] == N.HPCS_OK) and np.all(alone["status"] == N.HPCS_OK)
    want = np.array([_expected(oracle, r) for r in t], dtype=np.uint64)
    assert np.array_equal(mix["checksum"], want)
    assert np.array_equal(alone["checksum"], want)
