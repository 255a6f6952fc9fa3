"""Shared helpers for the tests: golden fixtures and closed-form checks."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

with open(os.path.join(HERE, "golden", "fixtures.json")) as _f:
    FIXTURES = json.load(_f)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def float_order(x: np.ndarray) -> np.ndarray:
    """Ascending by value with -0.0 before +0.0 (NaN-free input)."""
    return x[np.lexsort((~np.signbit(x), x))]


def hist_bytes(a: np.ndarray) -> np.ndarray:
    """Brute-force byte-digit counts via bincount."""
    P = a.dtype.itemsize
    rows = []
    for d in range(P):
        dg = (a >> np.array(8 * d, dtype=a.dtype)) & np.array(0xFF, dtype=a.dtype)
        rows.append(np.bincount(dg.astype(np.int64), minlength=256))
    return np.stack(rows).astype(np.int32)


def record_dtype(key_type=np.uint64):
    """C# sequential layout of struct { key; uint payload; } padded to the key's alignment."""
    ks = np.dtype(key_type).itemsize
    return np.dtype({"names": ["key", "payload"], "formats": [key_type, np.uint32],
                     "offsets": [0, ks], "itemsize": max(2 * ks, ks + 4) if ks == 8 else 8})
