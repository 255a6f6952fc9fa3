"""GPU parity of the serving loop: every request's result, run concurrently
with the others under continuous admission, has the checksum of the oracle's
result for the same payload, and the same checksum when the request runs
alone.  Exact (integer work, tolerance 0)."""
from __future__ import annotations

import numpy as np
import pytest

from hpcsharp_amd import _native as N
from hpcsharp_amd import serving as S
from tests._fixtures import float_order

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    yield


def _expected(oracle, req) -> int:
    x = S.payload(req)
    if int(req["op"]) == N.HPCS_OP_SORT:
        if x.dtype.kind == "f":
            oracle.sort_radix_float(x)
        else:
            oracle.sort_radix_lsd(x)
        return S.checksum(x)
    k = int(req["k"])
    if x.dtype == np.uint32:
        y = x.copy()
        if k < y.size:
            oracle.select_topk_u32(y, y.size - k)
        top = np.sort(y[y.size - k:])[::-1]
    elif x.dtype.kind == "f":
        top = float_order(x)[::-1][:k]
    else:
        top = np.sort(x)[::-1][:k]
    return S.checksum(np.ascontiguousarray(top))


def _trace():
    return S.make_trace(nreq=240, rate=20000.0, seed=0x5EED0042, n_lo=1 << 10, n_hi=1 << 20)


def test_results_alone_and_in_the_mix_match_the_oracle(oracle):
    t = _trace()
    mix, st = S.run(t, slots=4)
    alone, _ = S.run(t, mode=N.HPCS_SERVE_ALONE)
    assert np.all(mix["status"] == N.HPCS_OK) and np.all(alone["status"] == N.HPCS_OK)
    assert np.array_equal(mix["checksum"], alone["checksum"])
    want = np.array([_expected(oracle, r) for r in t], dtype=np.uint64)
    bad = np.nonzero(mix["checksum"] != want)[0]
    assert bad.size == 0, [(int(i), t[i]) for i in bad[:5]]
    # requests really overlapped: several slots busy, and some admitted while others ran
    assert len(set(mix["slot"].tolist())) == 4
    busy = [(a, d) for a, d in zip(mix["admit_s"], mix["done_s"])]
    overlaps = sum(1 for i, (a, d) in enumerate(busy) for (a2, d2) in busy[i + 1:] if a2 < d and a < d2)
    assert overlaps > 0
    s = S.summarize(t, mix, st)
    assert s["completed"] == t.size and s["p99_ms"] >= s["p50_ms"] > 0


def test_budget_admission_and_rejection():
    t = _trace()
    _, st = S.run(t, slots=2)                                 # the slots' scratch for this trace
    big = max(int(r["n"]) * S.DTYPES[int(r["dtype"])]().itemsize for r in t)
    budget = 2 * st.scratch_bytes_per_slot + 3 * big        # room for about three payloads at once
    res, st2 = S.run(t, slots=2, hbm_budget=budget)
    assert np.all(res["status"] == N.HPCS_OK)
    assert st2.peak_bytes <= st2.request_budget_bytes
    assert np.any(res["admit_s"] - t["arrival_s"] > 0)
    # a request larger than the whole request budget is rejected, the rest still run
    t2 = t.copy()
    t2["n"][5] = (1 << 29)
    t2["dtype"][5] = N.HPCS_DTYPE_U64
    t2["op"][5] = N.HPCS_OP_SORT
    res2, st3 = S.run(t2[:20], slots=2, hbm_budget=2 * st.scratch_bytes_per_slot + 3 * big)
    assert res2["status"][5] == N.HPCS_E_BUDGET
    assert np.all(np.delete(res2["status"], 5) == N.HPCS_OK)


def test_coalesced_small_sorts_match_the_oracle(oracle):
    """Small u32 / f32 sorts queued together run as one segmented sort: the
    same checksums as the oracle's results, alone or coalesced."""
    t = S.make_trace(nreq=400, rate=200000.0, seed=0x5EED0077, n_lo=1 << 4, n_hi=1 << 14, sort_frac=0.85,
                     dtypes=(N.HPCS_DTYPE_U32, N.HPCS_DTYPE_F32, N.HPCS_DTYPE_U64))
    res, st = S.run(t, slots=2, coalesce_max_keys=1 << 14, coalesce_max_requests=64)
    assert np.all(res["status"] == N.HPCS_OK)
    assert st.coalesced_requests > 50 and st.batches < t.size
    want = np.array([_expected(oracle, r) for r in t], dtype=np.uint64)
    bad = np.nonzero(res["checksum"] != want)[0]
    assert bad.size == 0, [(int(i), t[i]) for i in bad[:5]]
    alone, st1 = S.run(t, slots=2)
    assert st1.coalesced_requests == 0
    assert np.array_equal(alone["checksum"], res["checksum"])


def test_forced_deadline_fails_naming_a_request():
    """A deadline shorter than any request (forced from the host; every
    kernel still runs to its end) fails the run with HPCS_E_HIP and a
    message naming the request; a later run on the same device succeeds."""
    t = _trace()
    with pytest.raises(N.HpcsError) as e:
        S.run(t, slots=2, timeout_s=1e-7)
    assert e.value.code == N.HPCS_E_HIP
    assert "did not complete within" in str(e.value) and "request " in str(e.value)
    res, _ = S.run(t[:40], slots=2)
    assert np.all(res["status"] == N.HPCS_OK)
