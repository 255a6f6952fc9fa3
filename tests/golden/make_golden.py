"""Write tests/golden/fixtures.json: inputs and expected outputs for the sort path.

Expected outputs come from closed forms computed with numpy, never from the
oracle or the library under test:
  * keys-only sorts: the unique ascending order (np.sort);
  * stable key-value / record sorts: np.argsort(kind="stable");
  * float/double: numeric order with -0.0 before +0.0 (np.lexsort on
    (value, not signbit));
  * histograms: np.bincount per byte digit.
Large vectors are stored as SHA-256 digests of the little-endian bytes; the
inputs are regenerated from hpcsharp_amd.datagen (splitmix64, pinned by its
published known answers in the tests).  The reference's own example vectors
(HPCsharpExamples Program.cs:15-26 and SortOfUserDefinedClass.cs:86-103) are
stored in full.

Run: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from hpcsharp_amd import datagen  # noqa: E402


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def hist_bytes(a: np.ndarray) -> np.ndarray:
    P = a.dtype.itemsize
    return np.stack([np.bincount(((a >> np.array(8 * d, dtype=a.dtype)) & np.array(0xFF, dtype=a.dtype)).astype(np.int64),
                                 minlength=256) for d in range(P)]).astype(np.int32)


def float_order(x: np.ndarray) -> np.ndarray:
    return x[np.lexsort((~np.signbit(x), x))]


SIZES_U32 = [0, 1, 2, 3, 255, 256, 257, 4095, 4096, 4097, 65535, 65536, 65537, 100003]
DISTS = ["uniform", "low8", "sorted", "constant"]


def main():
    fx = {"generator": "splitmix64", "u32": [], "u64": [], "f64": [], "f32": [], "pairs": [], "hist": []}
    fx["splitmix64_seed0_first3"] = ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4", "0x06c45d188009454f"]
    # the reference's example vectors, in full
    fx["reference_examples"] = {
        "program_cs_uint": {"input": [21, 43, 16, 5, 54, 3], "expected": [3, 5, 16, 21, 43, 54]},
        "sorting_usage_long": {"input": [21, 43, 16, 5, 54, 3], "expected": [3, 5, 16, 21, 43, 54]},
        "user_defined_class": {
            "keys": [16, 12, 18, 18, 10, 2],
            "index": [0, 1, 2, 3, 4, 5],
            "expected_keys": [2, 10, 12, 16, 18, 18],
            "expected_index": [5, 4, 1, 0, 2, 3],
        },
    }
    seed = 0x5EED0100
    for dist in DISTS:
        for n in SIZES_U32:
            a = datagen.keys_u32(n, seed + n, dist)
            fx["u32"].append({"n": n, "dist": dist, "seed": seed + n, "input_sha": sha(a), "sorted_sha": sha(np.sort(a))})
    for n in [0, 1, 2, 257, 4097, 65537]:
        a = datagen.keys_u64(n, seed + 7 * n + 1)
        fx["u64"].append({"n": n, "seed": seed + 7 * n + 1, "input_sha": sha(a), "sorted_sha": sha(np.sort(a))})
    for n in [2, 1000, 65537]:
        x = datagen.doubles_special(n, seed + 11 * n)
        fx["f64"].append({"n": n, "seed": seed + 11 * n, "input_sha": sha(x), "sorted_sha": sha(float_order(x))})
        y = datagen.floats_finite(n, seed + 13 * n)
        fx["f32"].append({"n": n, "seed": seed + 13 * n, "input_sha": sha(y), "sorted_sha": sha(float_order(y))})
    specials = np.array([1.0, -0.0, 0.0, -1e300, 5e-324, -5e-324, 2.2250738585072014e-308, -2.2250738585072014e-308,
                         1e300, -1.0, float("inf"), float("-inf"), 0.0, -0.0, 2.225073858507201e-308], dtype=np.float64)
    fx["f64_specials"] = {"input_bits": [hex(v) for v in specials.view(np.uint64)],
                          "expected_bits": [hex(v) for v in float_order(specials).view(np.uint64)]}
    # stable pairs: narrow keys, payload = original index
    for n, mask in [(5000, 0xFF), (70000, 0xFFFF), (70000, 0xFFFFFFFFFFFFFFFF)]:
        k = datagen.keys_u64(n, seed + n + mask % 97) & np.uint64(mask)
        perm = np.argsort(k, kind="stable").astype(np.uint32)
        fx["pairs"].append({"n": n, "mask": hex(mask), "seed": seed + n + mask % 97, "input_sha": sha(k),
                            "sorted_keys_sha": sha(k[perm]), "perm_sha": sha(perm)})
    for n in [0, 1, 1000, 65537]:
        a = datagen.keys_u32(n, seed + 3 * n, "uniform")
        b = datagen.keys_u64(n, seed + 5 * n)
        fx["hist"].append({"n": n, "seed_u32": seed + 3 * n, "seed_u64": seed + 5 * n,
                           "u32_sha": sha(hist_bytes(a)), "u64_sha": sha(hist_bytes(b))})
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(fx, f, indent=1)
    print("wrote", os.path.join(HERE, "fixtures.json"))


if __name__ == "__main__":
    main()
