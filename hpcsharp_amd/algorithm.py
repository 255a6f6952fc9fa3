"""HPCsharp's radix-sort extension methods, served by the gfx950 library.

``Algorithm`` and ``ParallelAlgorithm`` keep the reference's names, argument
meaning, results and exceptions (namespace ``HPCsharp``, classes
``Algorithm`` in RadixSortLsd.cs:90 / Histogram.cs:30 and
``ParallelAlgorithm`` in RadixSortLsdParallel.cs:25 / HistogramParallel.cs:15).
Arrays are numpy arrays standing in for the managed arrays: sorts run in
place on the caller's array and the ``...Func`` forms return that same array,
as the reference does.  Every call goes through libhpcsort.so on the GPU;
there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as N


class ArgumentException(ValueError):
    """System.ArgumentException (e.g. PartitionRadix.cs:180-181); as in .NET,
    the two below derive from it."""


class ArgumentNullException(ArgumentException):
    """System.ArgumentNullException (e.g. RadixSortLsd.cs:215-216)."""


class ArgumentOutOfRangeException(ArgumentException):
    """System.ArgumentOutOfRangeException (e.g. RadixSortLsd.cs:1621-1622)."""


class OutOfMemoryException(MemoryError):
    """System.OutOfMemoryException."""


class InvalidOperationException(RuntimeError):
    """A HIP runtime failure, carrying the library's message."""


def _raise(err: N.HpcsError):
    if err.code == N.HPCS_E_INVALID_ARG:
        raise ArgumentOutOfRangeException(str(err)) from None
    if err.code in (N.HPCS_E_OOM, N.HPCS_E_BUDGET):
        raise OutOfMemoryException(str(err)) from None
    raise InvalidOperationException(str(err)) from None


def _call(fn, *args, ctx=None):
    try:
        N.check(fn(N.context() if ctx is None else ctx, *args))
    except N.HpcsError as e:
        _raise(e)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _require(a, name: str, dtypes):
    if a is None:
        raise ArgumentNullException(name)
    if not isinstance(a, np.ndarray) or a.ndim != 1 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError(f"{name} must be a contiguous one-dimensional numpy array")
    if a.dtype not in dtypes:
        raise TypeError(f"{name} has unsupported element type {a.dtype}")
    if not a.flags["WRITEABLE"]:
        raise TypeError(f"{name} is read-only")


_SORT = {
    np.dtype(np.uint32): "hpcs_sort_radix_u32",
    np.dtype(np.uint64): "hpcs_sort_radix_u64",
    np.dtype(np.float32): "hpcs_sort_radix_f32",
    np.dtype(np.float64): "hpcs_sort_radix_f64",
    np.dtype(np.int32): "hpcs_sort_radix_i32",
    np.dtype(np.int64): "hpcs_sort_radix_i64",
}


_DIGIT_CTX: dict[int, ctypes.c_void_p] = {}


def _digit_context(bits: int):
    """A context whose sorts use `bits`-bit digits where the library has that
    form (11: the single-pass form of keys without values), created once."""
    ctx = _DIGIT_CTX.get(bits)
    if ctx is None:
        ctx = N.new_context(0)
        N.set_option(N.HPCS_OPT_DIGIT_BITS, bits, ctx)
        _DIGIT_CTX[bits] = ctx
    return ctx


def _sort_whole(a: np.ndarray, ctx=None) -> None:
    _call(getattr(N.lib(), _SORT[a.dtype]), _ptr(a), a.size, ctx=ctx)


def _sort_range(a: np.ndarray, startIndex: int, length: int, ctx=None) -> None:
    if startIndex < 0 or length < 0 or startIndex + length > a.size:
        raise ArgumentOutOfRangeException("startIndex/length outside the array")
    if a.dtype == np.uint32:
        _call(N.lib().hpcs_sort_radix_range_u32, _ptr(a), startIndex, length, ctx=ctx)
    else:
        sub = a[startIndex:startIndex + length]
        _call(getattr(N.lib(), _SORT[a.dtype]), _ptr(sub), sub.size, ctx=ctx)


def _extract_keys(items, getKey, start: int, length: int, keyBits: int | None = None) -> np.ndarray:
    """getKey (the reference's Func<T, UInt32 / UInt64>) applied on the host
    to items[start, start + length): uint32 keys for keyBits 32 (a UInt32
    extractor), uint64 for 64, and without keyBits uint32 when every key fits
    32 bits (the sorted order is the same either way).  A key the width does
    not hold raises ArgumentOutOfRangeException."""
    if keyBits not in (None, 32, 64):
        raise ArgumentOutOfRangeException("keyBits must be 32 or 64")
    vals = [int(getKey(items[start + i])) for i in range(length)]
    limit = 1 << (keyBits or 64)
    if any(v < 0 or v >= limit for v in vals):
        raise ArgumentOutOfRangeException(f"getKey returned a value outside UInt{keyBits or 64}")
    wide = keyBits == 64 or (keyBits is None and any(v >= 1 << 32 for v in vals))
    return np.array(vals, dtype=np.uint64 if wide else np.uint32)


def _order_by_keys(keys: np.ndarray) -> np.ndarray:
    """The stable order of keys: (key, index) pairs sorted on the GPU, the
    indices returned (keys is sorted in place too)."""
    idx = np.arange(keys.size, dtype=np.uint32)
    fn = N.lib().hpcs_sort_radix_pairs_u64_u32 if keys.dtype == np.uint64 else N.lib().hpcs_sort_radix_pairs_u32_u32
    _call(fn, _ptr(keys), _ptr(idx), keys.size)
    return idx


def _sort_by_extractor(items, getKey, start: int = 0, length: int | None = None):
    """Stable sort of items[start, start + length) by getKey, in place: the
    keys are extracted on the host (a delegate cannot run on the GPU), the
    (key, index) pairs sorted on the GPU, the items gathered on the host."""
    if items is None:
        raise ArgumentNullException("inputArray")
    if getKey is None:
        raise ArgumentNullException("getKey")
    if length is None:
        length = len(items) - start
    if start < 0 or length < 0 or start + length > len(items):
        raise ArgumentOutOfRangeException("start/length outside the array")
    idx = _order_by_keys(_extract_keys(items, getKey, start, length)).astype(np.int64) + start
    if isinstance(items, np.ndarray):
        items[start:start + length] = items[idx]
    else:
        items[start:start + length] = [items[i] for i in idx]
    return items


class Algorithm:
    """Serial-surface methods (namespace HPCsharp, class Algorithm).

    Key-extractor overloads take getKey as a Python callable (the reference's
    Func<T, UInt32 / UInt64>) over a numpy array (the reference's T[]: sorted
    in place and returned) or, where the reference has a List<T> overload, a
    Python list (a new sorted list is returned, as the reference returns a new
    List<T>)."""

    # -- uint[] / ulong[] LSD --------------------------------------------------
    @staticmethod
    def SortRadixLsd(inOutArray, startIndex=None, length: int | None = None):
        """SortRadixLsd(this uint[]) RadixSortLsd.cs:255-260; (start, length) :242-247
        with the Experimental overload's sub-range semantics (:1409-1457);
        ulong[] :424-442; int[] / long[] :267-357, 444-532 (signed order).
        SortRadixLsd(List<uint>) :615-623: a Python list of keys gives a new
        sorted list.  SortRadixLsd<T>(this T[], Func<T, UInt32 / UInt64>)
        :714-750, :783-820: a callable second argument sorts the items stably
        by it (all 64 key bits; the reference's UInt64 form stops after four
        bytes, :795, :814) and returns them."""
        if callable(startIndex):
            return _sort_by_extractor(inOutArray, startIndex)
        if isinstance(inOutArray, list):
            a = np.array(inOutArray, dtype=np.uint32)
            _sort_whole(a)
            return [int(x) for x in a]
        _require(inOutArray, "inOutArray", (np.uint32, np.uint64, np.int32, np.int64))
        if startIndex is None:
            _sort_whole(inOutArray)
        else:
            _sort_range(inOutArray, startIndex, inOutArray.size - startIndex if length is None else length)

    @staticmethod
    def SortRadixLsd11bit(inOutArray: np.ndarray, startIndex: int | None = None,
                          length: int | None = None) -> np.ndarray:
        """SortRadixLsd11bit(this uint[]) RadixSortLsd.cs:1540-1546 and
        (startIndex, length) :1484-1532: sorted with 11-bit digits (three
        passes) where the library's single-pass form serves the size, with
        8-bit digits otherwise (a stable LSD sort's output does not depend on
        the digit width).  Returns the array, sorted; the reference's
        whole-array overload returns its pass-2 state (odd pass count,
        :1540-1546), a quirk not reproduced."""
        _require(inOutArray, "inOutArray", (np.uint32,))
        ctx = _digit_context(11)
        if startIndex is None:
            _sort_whole(inOutArray, ctx)
        else:
            _sort_range(inOutArray, startIndex, inOutArray.size - startIndex if length is None else length, ctx)
        return inOutArray

    @staticmethod
    def SortRadixLsdNbit(inOutArray: np.ndarray, startIndex: int, length: int, bitsPerDigit: int) -> None:
        """SortRadixLsdNbit(this uint[] / ulong[], startIndex, length,
        bitsPerDigit) RadixSortLsd.cs:1616-1661, 1839-1882:
        ArgumentOutOfRangeException when bitsPerDigit exceeds the key width
        (:1621-1622, :1844-1845) or is below 1.  The range is sorted in place
        (the Experimental overload's sub-range semantics); 11-bit digits are
        used as in SortRadixLsd11bit, any other width sorts with the library's
        8-bit passes (same output)."""
        _require(inOutArray, "inOutArray", (np.uint32, np.uint64))
        if bitsPerDigit < 1 or bitsPerDigit > 8 * inOutArray.itemsize:
            raise ArgumentOutOfRangeException("bitsPerDigit")
        _sort_range(inOutArray, startIndex, length, _digit_context(11) if bitsPerDigit == 11 else None)

    @staticmethod
    def SortRadixLsdFunc(inOutArray: np.ndarray, startIndex: int | None = None, length: int | None = None) -> np.ndarray:
        """SortRadixLsdFunc(this uint[]) RadixSortLsd.cs:213-232 (returns the same array)."""
        Algorithm.SortRadixLsd(inOutArray, startIndex, length)
        return inOutArray

    @staticmethod
    def SortRadix(inputArray, keys=None, length: int | None = None, outputArray=None, getKey=None):
        """The reference shapes that share this name:
        - SortRadix(this uint[]) -- NuGet 3.17.0's name for SortRadixLsdFunc
          (HPCsharpExamples Program.cs:26): sorts in place, returns the array;
        - SortRadix<T>(T[] items, UInt32[]/UInt64[] keys) RadixSortLsd.cs:
          1183-1299: items (any fixed-size dtype, e.g. a structured array)
          sorted stably by a parallel key array, returned as (items, keys);
          with a list of items, SortRadix<T>(List<T>, keys) :1307-1338: a new
          list and the sorted keys;
        - SortRadix<T>(List<T>, Func<T, UInt32 / UInt64>) :1151-1182: a new
          list sorted stably by the callable;
        - SortRadix<T>(T[] in, start, length, T[] out, Func<T, UInt32>)
          :1091-1135, called SortRadix(a, start, length, out, getKey): the
          range sorted in `a` and copied to `out` (the reference leaves it in
          both arrays)."""
        if getKey is not None or outputArray is not None:
            start = keys
            if inputArray is None or outputArray is None:
                raise ArgumentNullException("inputArray" if inputArray is None else "outputArray")
            if start is None or length is None:
                raise ArgumentOutOfRangeException("start/length")
            if start < 0 or length < 0 or start + length > min(len(inputArray), len(outputArray)):
                raise ArgumentOutOfRangeException("start/length outside the arrays")
            _sort_by_extractor(inputArray, getKey, start, length)
            outputArray[start:start + length] = inputArray[start:start + length]
            return None
        if callable(keys):
            if inputArray is None:
                raise ArgumentNullException("inputList")
            return _sort_by_extractor(list(inputArray), keys)
        if isinstance(inputArray, list) and keys is not None:
            _require(keys, "inKeys", (np.uint32, np.uint64))
            if keys.size != len(inputArray):
                raise ArgumentOutOfRangeException("inKeys must have one key per item")
            idx = _order_by_keys(keys)
            return [inputArray[i] for i in idx], keys
        if keys is None:
            _require(inputArray, "inputArray", tuple(_SORT))
            _sort_whole(inputArray)
            return inputArray
        if inputArray is None:
            raise ArgumentNullException("inputArray")
        _require(keys, "inKeys", (np.uint32, np.uint64))
        if not isinstance(inputArray, np.ndarray) or inputArray.ndim != 1 or not inputArray.flags["C_CONTIGUOUS"]:
            raise TypeError("inputArray must be a contiguous one-dimensional numpy array")
        if keys.size != inputArray.size:
            raise ArgumentOutOfRangeException("inKeys must have one key per item")
        if inputArray.dtype == np.uint32:
            fn = N.lib().hpcs_sort_radix_pairs_u64_u32 if keys.dtype == np.uint64 else N.lib().hpcs_sort_radix_pairs_u32_u32
            _call(fn, _ptr(keys), _ptr(inputArray), keys.size)
        else:                                   # items of any fixed size
            _call(N.lib().hpcs_sort_radix_keys_items, _ptr(keys), 8 * keys.dtype.itemsize, _ptr(inputArray),
                  inputArray.dtype.itemsize, keys.size)
        return inputArray, keys

    # -- MSD entry points ------------------------------------------------------
    @staticmethod
    def SortRadixMsd(arrayToBeSorted: np.ndarray, start: int | None = None, length: int | None = None,
                     threshold: int = 1024) -> None:
        """SortRadixMsd(this uint[] [, start, length] [, threshold]) RadixSortMsd.cs:682-712,
        ulong[] :813-838, int[] :1285, long[] :1308, float[]/double[] :1563-1595 (ascending by
        the radix key, -0.0 before +0.0).  The reference sorts in place MSD-first with an
        Array.Sort base case below `threshold` (accepted here, unused); the result of a
        sort is the sorted array whatever the algorithm, and the library's stable LSD
        passes produce it (with n scratch keys, unlike the in-place MSD form)."""
        _require(arrayToBeSorted, "arrayToBeSorted",
                 (np.uint32, np.uint64, np.int32, np.int64, np.float32, np.float64))
        if start is None:
            _sort_whole(arrayToBeSorted)
        else:
            _sort_range(arrayToBeSorted, start, arrayToBeSorted.size - start if length is None else length)

    @staticmethod
    def PartitionRadix(arrayToBePartitioned: np.ndarray, *args, threshold: int = 1024) -> None:
        """PartitionRadix(this uint[], [start, length,] k | int[] k [, threshold])
        PartitionRadix.cs:135-210: afterwards a[k] holds the k-th smallest key of the
        range, no key before it is larger and none after it smaller, for every k given
        (the reference's multi-select leaves the rest in no particular order; a sorted
        range meets the same contract, and is what this returns).  Argument checks as
        the reference's: a null array or k array -> ArgumentNullException, an empty
        array or range -> ArgumentOutOfRangeException, more indices than keys ->
        ArgumentException, no indices -> no change."""
        a = arrayToBePartitioned
        if a is None:
            raise ArgumentNullException("arrayToBePartitioned")
        if len(args) == 3:
            start, length, k = args
        elif len(args) == 1:
            start, length, k = 0, None, args[0]
        else:
            raise TypeError("PartitionRadix(a, k) or PartitionRadix(a, start, length, k)")
        _require(a, "arrayToBePartitioned", (np.uint32,))
        if k is None:
            raise ArgumentNullException("k")
        if length is None:
            length = a.size
        if a.size <= 0 or length <= 0:
            raise ArgumentOutOfRangeException("array length or length is invalid")
        ks = [k] if isinstance(k, (int, np.integer)) else list(k)
        if len(ks) > a.size:
            raise ArgumentException("k array length cannot be greater than the array to be selected length")
        if not ks:
            return
        _sort_range(a, start, length)

    @staticmethod
    def SortRadixMsdInPlaceFunc(arrayToBeSorted: np.ndarray) -> np.ndarray:
        """RadixSortMsd.cs:1578,1602."""
        Algorithm.SortRadixMsd(arrayToBeSorted)
        return arrayToBeSorted

    # -- key-extractor records -------------------------------------------------
    @staticmethod
    def SortRadixLsd2(inputArray, getKey):
        """SortRadixLsd2<T>(this T[], Func<T,UInt32/UInt64>) RadixSortLsd.cs:751-775, 821-846.

        With a callable getKey (the Func), the items (a numpy array of any
        dtype, or a list) are sorted stably in place by it and returned.  With
        a field name of a numpy structured array (the unmanaged struct form;
        uint32/uint64/float32/float64 key), the GPU reads the key from each
        record and moves the records."""
        if inputArray is None:
            raise ArgumentNullException("inputArray")
        if getKey is None:
            raise ArgumentNullException("getKey")
        if callable(getKey):
            return _sort_by_extractor(inputArray, getKey)
        if not isinstance(inputArray, np.ndarray) or inputArray.dtype.fields is None:
            raise TypeError("inputArray must be a numpy structured array")
        if not inputArray.flags["C_CONTIGUOUS"]:
            raise TypeError("inputArray must be contiguous")
        ftype, offset = inputArray.dtype.fields[getKey][:2]
        kinds = {np.dtype(np.uint32): (32, N.HPCS_KEY_UINT), np.dtype(np.uint64): (64, N.HPCS_KEY_UINT),
                 np.dtype(np.float32): (32, N.HPCS_KEY_FLOAT), np.dtype(np.float64): (64, N.HPCS_KEY_FLOAT)}
        if ftype not in kinds:
            raise TypeError(f"key field {getKey} has unsupported type {ftype}")
        bits, kind = kinds[ftype]
        _call(N.lib().hpcs_sort_radix_records, _ptr(inputArray), inputArray.size, inputArray.dtype.itemsize,
              offset, bits, kind)
        return inputArray

    SortRadixLsd_T = SortRadixLsd2
    SortRadixFaster2 = SortRadixLsd2

    @staticmethod
    def SortRadixFaster(inputArray, getKey):
        """SortRadixFaster<T>(this T[], Func<T, UInt32 / UInt64>) RadixSortLsd.cs:
        861-915, 961-1014 (all 64 key bits; the reference's UInt64 form stops
        after four bytes, :975, :1006): as SortRadixLsd2.  With a list,
        SortRadixFaster<T>(List<T>, Func) :1059-1090: a new sorted list."""
        if isinstance(inputArray, list):
            if getKey is None:
                raise ArgumentNullException("getKey")
            return _sort_by_extractor(list(inputArray), getKey)
        return Algorithm.SortRadixLsd2(inputArray, getKey)

    # -- histograms ------------------------------------------------------------
    @staticmethod
    def HistogramByteComponents(inArray, l: int, r: int, getKey=None, keyBits: int = 32) -> np.ndarray:
        """HistogramByteComponents(uint[]/ulong[], int l, int r) Histogram.cs:261-296, 957-990:
        int32[digits][256] of [l, r] inclusive; int[] / long[] :992-1052 count
        the top byte ^ 128.  With getKey (a field name of a structured array,
        the key extractor Func<T, UInt32/UInt64>, Histogram.cs:760-844) the
        bytes of each record's key are counted."""
        if inArray is None:
            raise ArgumentNullException("inArray")
        if callable(getKey):
            return Algorithm.HistogramByteComponentsAndKeyArray(inArray, l, r, getKey, keyBits)[0]
        if getKey is not None:
            if not isinstance(inArray, np.ndarray) or inArray.dtype.names is None or getKey not in inArray.dtype.names:
                raise TypeError("getKey must name a field of a structured array")
            kt, off = inArray.dtype.fields[getKey][:2]
            if kt not in (np.dtype(np.uint32), np.dtype(np.uint64)):
                raise TypeError("the key field must be uint32 or uint64")
            if l < 0 or (r >= l and r >= inArray.size):
                raise IndexError("l..r outside the array")
            out = np.zeros((kt.itemsize, 256), dtype=np.int32)
            _call(N.lib().hpcs_histogram_byte_components_records, _ptr(np.ascontiguousarray(inArray)), l, r,
                  inArray.dtype.itemsize, off, 8 * kt.itemsize, _ptr(out))
            return out
        _require(inArray, "inArray", (np.uint32, np.uint64, np.int32, np.int64))
        if l < 0 or (r >= l and r >= inArray.size):
            raise IndexError("l..r outside the array")
        P = inArray.dtype.itemsize
        out = np.zeros((P, 256), dtype=np.int32)
        fn = getattr(N.lib(), "hpcs_histogram_byte_components_" + ("i" if inArray.dtype.kind == "i" else "u") + str(8 * P))
        _call(fn, _ptr(inArray), l, r, _ptr(out))
        return out

    @staticmethod
    def HistogramByteComponentsAndKeyArray(inArray, l: int, r: int, getKey, keyBits: int = 32):
        """HistogramByteComponentsAndKeyArray<T>(T[], l, r, Func<T, UInt32 /
        UInt64>) Histogram.cs:846-934 (and HistogramByteComponents<T> with a
        Func, :760-844): the callable's keys of [l, r] extracted on the host,
        their byte components counted on the GPU.  Returns (int32[digits][256],
        keys), keys holding each key at its index (0 outside [l, r]); all 8
        bytes of 64-bit keys are counted (the reference's UInt64 form sizes 4
        rows and throws, :894, :904).  keyBits (32 or 64) is the extractor's
        result type, UInt32 or UInt64."""
        if inArray is None:
            raise ArgumentNullException("inArray")
        if getKey is None:
            raise ArgumentNullException("getKey")
        if l < 0 or (r >= l and r >= len(inArray)):
            raise IndexError("l..r outside the array")
        span = max(0, r - l + 1)
        k = _extract_keys(inArray, getKey, l, span, keyBits)
        keys = np.zeros(len(inArray), dtype=k.dtype)
        keys[l:l + span] = k
        return Algorithm.HistogramByteComponents(keys, l, r), keys

    @staticmethod
    def Histogram11bitComponents(inArray: np.ndarray, l: int, r: int) -> np.ndarray:
        """Histogram11bitComponents(uint[], l, r) Histogram.cs:298-331: int32[3][2048]."""
        return Algorithm.HistogramNbitComponents(inArray, l, r, 11)

    @staticmethod
    def HistogramNbitComponents(inArray: np.ndarray, l: int, r: int, numberOfBitPerComponent: int) -> np.ndarray:
        """HistogramNbitComponents(uint[]/ulong[], l, r, bits) Histogram.cs:381-427:
        int32[ceil(width / bits)][2^bits] (bits in [1, 16] here)."""
        if inArray is None:
            raise ArgumentNullException("inArray")
        _require(inArray, "inArray", (np.uint32, np.uint64))
        if l < 0 or (r >= l and r >= inArray.size):
            raise IndexError("l..r outside the array")
        bits = int(numberOfBitPerComponent)
        if not 1 <= bits <= 16:
            raise ArgumentOutOfRangeException("numberOfBitPerComponent")
        width = 8 * inArray.dtype.itemsize
        out = np.zeros(((width + bits - 1) // bits, 1 << bits), dtype=np.int32)
        fn = N.lib().hpcs_histogram_nbit_components_u32 if width == 32 else N.lib().hpcs_histogram_nbit_components_u64
        _call(fn, _ptr(inArray), l, r, bits, _ptr(out))
        return out

    # -- radix top-k -------------------------------------------------------------
    @staticmethod
    def SelectTopK(arrayToBeSelected: np.ndarray, start: int, length: int | None = None, k: int | None = None) -> None:
        """SelectTopK(this uint[] a, int k) SelectionTopKRadix.cs:160-171 and
        SelectTopK(this uint[] a, int start, int length, int k) :142-153.

        Called as SelectTopK(a, k) or SelectTopK(a, start, length, k).  k is an
        index: afterwards a[k:start+length] holds the largest keys of the range in
        ascending order (a[k] is the k-th smallest) and a[start:k] the others."""
        if arrayToBeSelected is None:
            raise ArgumentNullException("arrayToBeSelected")
        _require(arrayToBeSelected, "arrayToBeSelected", (np.uint32,))
        if length is None and k is None:
            start, length, k = 0, arrayToBeSelected.size, start
            if length <= 0:
                raise ArgumentOutOfRangeException("array length is invalid")
        if start < 0 or length is None or length <= 0 or start + length > arrayToBeSelected.size:
            raise ArgumentOutOfRangeException("start or length are invalid")
        if k is None or k < start or k >= start + length:
            raise ArgumentOutOfRangeException("k must be between start and (start + length)")
        _call(N.lib().hpcs_select_topk_u32, _ptr(arrayToBeSelected), start, length, k)

    @staticmethod
    def TopK(inputArray: np.ndarray, k: int) -> np.ndarray:
        """The k largest keys of a uint[]/ulong[]/float[]/double[] array in
        descending order (the radix order of the sort: -0.0 below +0.0); the
        input is not modified."""
        _require(inputArray, "inputArray", tuple(_TOPK))
        if k < 0 or k > inputArray.size:
            raise ArgumentOutOfRangeException("k must lie in [0, n]")
        out = np.empty(k, dtype=inputArray.dtype)
        _call(getattr(N.lib(), _TOPK[inputArray.dtype]), _ptr(inputArray), inputArray.size, k, _ptr(out))
        return out


_TOPK = {
    np.dtype(np.uint32): "hpcs_topk_u32",
    np.dtype(np.uint64): "hpcs_topk_u64",
    np.dtype(np.float32): "hpcs_topk_f32",
    np.dtype(np.float64): "hpcs_topk_f64",
    np.dtype(np.int32): "hpcs_topk_i32",
    np.dtype(np.int64): "hpcs_topk_i64",
}

class ParallelAlgorithm:
    """Multicore-surface methods (namespace HPCsharp, class ParallelAlgorithm)."""

    @staticmethod
    def SortRadixPar(inputArray: np.ndarray, ParallelWorkQuantum: int = 64 * 1024) -> np.ndarray:
        """SortRadixPar(this uint[], int ParallelWorkQuantum = 64Ki) RadixSortLsdParallel.cs:145-232,
        (this ulong[]) :690-734 and the int[] / long[] forms :446-668.  The work
        quantum is accepted and unused: the GPU tiles are fixed by the kernel."""
        if inputArray is None:
            raise ArgumentNullException("inputArray")
        _require(inputArray, "inputArray", (np.uint32, np.uint64, np.int32, np.int64))
        _sort_whole(inputArray)
        return inputArray

    @staticmethod
    def HistogramByteComponentsPar(inArray: np.ndarray, l: int, r: int, parallelThreshold: int = 16 * 1024) -> np.ndarray:
        """HistogramByteComponentsPar(uint[]/ulong[], l, r, threshold) HistogramParallel.cs:214-222,
        545-555: the same counts as uint32[digits][256]."""
        return Algorithm.HistogramByteComponents(inArray, l, r).astype(np.uint32)
