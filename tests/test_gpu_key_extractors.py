"""GPU parity of the key-extractor (delegate) overloads in the Python mirror:
a callable getKey (the reference's Func<T, UInt32 / UInt64>) runs on the
host, the (key, index) pairs are sorted on the GPU, the items gathered.

Checked against the oracle's restatement of SortRadixLsd2 over records
(RadixSortLsd.cs:821-846), against Python's stable `sorted` for objects of a
user-defined class (the reference's SortOfUserDefinedClass.cs:86-103 example
included), and the List<T> overloads against their contract (a new sorted
list, the input unchanged)."""
from __future__ import annotations

import numpy as np
import pytest

from hpcsharp_amd import datagen
from hpcsharp_amd.algorithm import (Algorithm, ArgumentNullException, ArgumentOutOfRangeException,
                                    ParallelAlgorithm)

pytestmark = pytest.mark.gpu

REC = np.dtype({"names": ["key", "payload"], "formats": [np.uint64, np.uint32], "offsets": [0, 8], "itemsize": 16})


def _same(a: np.ndarray, b: np.ndarray) -> bool:
    """Records equal field by field (numpy leaves the 4 padding bytes of a
    copied record uninitialised, so raw bytes are not compared)."""
    return bool(np.array_equal(a["key"], b["key"]) and np.array_equal(a["payload"], b["payload"]))


def _records(n: int, seed: int, mask: int = 0xFFFFFFFFFFFFFFFF) -> np.ndarray:
    r = np.zeros(n, dtype=REC)
    r["key"] = datagen.keys_u64(n, seed) & np.uint64(mask)
    r["payload"] = np.arange(n, dtype=np.uint32)
    return r


class UserRecord:
    """A user-defined class (reference objects: nothing the GPU can move)."""

    def __init__(self, key: int, index: int):
        self.key = key
        self.index = index


@pytest.mark.parametrize("mask", [0xFFFFFFFFFFFFFFFF, 0xFFFF, 0xFF00000000000000])
@pytest.mark.parametrize("method", ["SortRadixLsd2", "SortRadixLsd", "SortRadixFaster", "SortRadixFaster2"])
def test_callable_key_matches_the_oracle_record_sort(oracle, method, mask):
    r = _records(30011, 0x5EED0007 + mask % 97, mask)
    want = r.copy()
    oracle.sort_radix_records(want, 0, 64)
    got = getattr(Algorithm, method)(r, lambda e: e["key"])
    assert got is r
    assert _same(r, want), (method, hex(mask))


def test_callable_and_field_name_forms_agree(oracle):
    r = _records(50003, 42, 0xFFFFF)
    a, b = r.copy(), r.copy()
    Algorithm.SortRadixLsd2(a, lambda e: e["key"])
    Algorithm.SortRadixLsd2(b, "key")
    assert _same(a, b)


def test_reference_example_user_defined_class():
    """HPCsharpExamples SortOfUserDefinedClass.cs:86-103: keys {16, 12, 18,
    18, 10, 2} with their indices sort stably to (2,5) (10,4) (12,1) (16,0)
    (18,2) (18,3)."""
    items = [UserRecord(k, i) for i, k in enumerate([16, 12, 18, 18, 10, 2])]
    out = Algorithm.SortRadixLsd(items, lambda u: u.key)
    assert [(u.key, u.index) for u in out] == [(2, 5), (10, 4), (12, 1), (16, 0), (18, 2), (18, 3)]


@pytest.mark.parametrize("bits", [32, 64])
def test_objects_sorted_stably(bits):
    rng = np.random.default_rng(bits)
    n = 20000
    keys = rng.integers(0, 1 << (bits - 8), n, dtype=np.uint64) & np.uint64(0xFF00FF if bits == 32 else 0xFF0000FF00FF)
    items = [UserRecord(int(k), i) for i, k in enumerate(keys)]
    want = sorted(items, key=lambda u: u.key)                  # Python's sort is stable
    got = Algorithm.SortRadixLsd2(list(items), lambda u: u.key)
    assert [u.index for u in got] == [u.index for u in want]


def test_list_overloads_return_a_new_list():
    items = [UserRecord(int(k), i) for i, k in enumerate(datagen.keys_u32(5000, 9) & np.uint32(0xFFF))]
    before = [u.index for u in items]
    want = [u.index for u in sorted(items, key=lambda u: u.key)]
    for fn in (Algorithm.SortRadix, Algorithm.SortRadixFaster):
        out = fn(items, lambda u: u.key)
        assert out is not items and [u.index for u in out] == want
        assert [u.index for u in items] == before               # the input List is unchanged
    keys = np.array([u.key for u in items], dtype=np.uint32)
    out, k = Algorithm.SortRadix(items, keys)                   # SortRadix<T>(List<T>, UInt32[])
    assert [u.index for u in out] == want and np.array_equal(k, np.sort(keys))
    lst = [int(x) for x in datagen.keys_u32(3001, 3)]           # SortRadixLsd(List<uint>)
    assert Algorithm.SortRadixLsd(lst) == sorted(lst)


def test_range_form_with_output_array():
    """SortRadix<T>(T[] in, start, length, T[] out, Func<T, UInt32>): the range
    sorted in `in` and copied to `out`; everything outside it untouched."""
    r = _records(10007, 5, 0xFFFFFFFF)
    out = np.zeros_like(r)
    src = r.copy()
    Algorithm.SortRadix(r, 1000, 7000, out, lambda e: e["key"])
    seg = src[1000:8000]
    want = seg[np.argsort(seg["key"], kind="stable")]
    assert _same(r[1000:8000], want)
    assert _same(out[1000:8000], want)
    assert _same(r[:1000], src[:1000])
    assert _same(r[8000:], src[8000:])
    assert not out[:1000]["key"].any() and not out[8000:]["key"].any()
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.SortRadix(r, 9000, 2000, out, lambda e: e["key"])


@pytest.mark.parametrize("bits", [32, 64])
def test_histogram_with_key_extractor(bits):
    r = _records(40009, 77, 0xFFFFFFFF if bits == 32 else 0xFFFFFFFFFFFFFFFF)
    l, rr = 17, 40000
    h, keys = Algorithm.HistogramByteComponentsAndKeyArray(r, l, rr, lambda e: e["key"], keyBits=bits)
    k = r["key"][l:rr + 1]
    want = np.stack([np.bincount((k >> np.uint64(8 * d)) & np.uint64(255), minlength=256)
                     for d in range(bits // 8)])
    assert h.shape == (bits // 8, 256) and np.array_equal(h, want)
    assert np.array_equal(keys[l:rr + 1], k) and not keys[:l].any() and not keys[rr + 1:].any()
    assert np.array_equal(Algorithm.HistogramByteComponents(r, l, rr, lambda e: e["key"], bits), want)


def test_generic_items_by_key_array():
    """SortRadix<T>(T[] items, UInt64[] keys) over records: stable by the full
    64-bit key (the reference's form stops after four passes, :1258)."""
    r = _records(20011, 21)
    keys = (datagen.keys_u64(r.size, 22) & np.uint64(0xFF000000000000FF)).copy()
    order = np.argsort(keys, kind="stable")
    want = r[order]
    items, k = Algorithm.SortRadix(r, keys)
    assert _same(items, want) and np.array_equal(k, np.sort(keys))


def test_extractor_argument_checks():
    with pytest.raises(ArgumentNullException):
        Algorithm.SortRadixLsd2(None, lambda e: 0)
    with pytest.raises(ArgumentNullException):
        Algorithm.SortRadixFaster([1, 2], None)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.SortRadixLsd2([UserRecord(-1, 0)], lambda u: u.key)
    with pytest.raises(ArgumentOutOfRangeException):
        Algorithm.HistogramByteComponents(_records(10, 1), 0, 9, lambda e: e["key"] | (1 << 40), keyBits=32)
    assert ParallelAlgorithm.SortRadixPar(np.zeros(0, np.int32)).size == 0
