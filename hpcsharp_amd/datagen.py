"""Seeded synthetic inputs for the sort workloads (host side, numpy).

x_i is output i of splitmix64 seeded with `seed` (Steele, Lea & Flood 2014):
z = seed + (i + 1) * 0x9E3779B97F4A7C15, then the 30/27/31 xor-shift-multiply
finaliser.  Integer-only, so every run and machine reproduces the arrays bit
for bit.

Distributions of the single-array u32 workload (the reference's harness
toggles the same three plus random, HPCsharpExamples SortSpeedup.cs:22-24):
  uniform  -- (uint32) x_i over the full range
  low8     -- C | (x_i & 0xFF), C a seed-derived constant with a zero low byte
  sorted   -- key = i (the reference's presorted generator)
  constant -- key = 0 (the reference's constant generator)
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, start: int = 0, chunk: int = 1 << 24) -> np.ndarray:
    """uint64 array of outputs [start, start + n) of splitmix64(seed)."""
    out = np.empty(n, dtype=np.uint64)
    s = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        for lo in range(0, n, chunk):
            hi = min(n, lo + chunk)
            z = np.arange(start + lo + 1, start + hi + 1, dtype=np.uint64)
            z *= _GOLDEN
            z += s
            z ^= z >> np.uint64(30)
            z *= _M1
            z ^= z >> np.uint64(27)
            z *= _M2
            z ^= z >> np.uint64(31)
            out[lo:hi] = z
    return out


def keys_u32(n: int, seed: int, dist: str = "uniform") -> np.ndarray:
    if dist == "uniform":
        return splitmix64(seed, n).astype(np.uint32)
    if dist == "low8":
        c = np.uint32(int(splitmix64(seed ^ 0xC0FFEE, 1)[0]) & 0xFFFFFF00)
        return (splitmix64(seed, n).astype(np.uint32) & np.uint32(0xFF)) | c
    if dist == "sorted":
        return np.arange(n, dtype=np.uint32)
    if dist == "constant":
        return np.zeros(n, dtype=np.uint32)
    raise ValueError(f"unknown distribution {dist}")


def keys_u64(n: int, seed: int) -> np.ndarray:
    return splitmix64(seed, n)


def uniform01(n: int, seed: int) -> np.ndarray:
    """Doubles in (0, 1) from the top 53 bits."""
    return ((splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / (1 << 53))


def doubles_special(n: int, seed: int, sigma: float = 1e3) -> np.ndarray:
    """N(0, sigma) by Box-Muller on splitmix64 uniforms, then 1% +-0.0, 1%
    subnormals (random mantissa and sign) and 1% +-1e300 at disjoint
    seed-chosen positions (the floating-key workload)."""
    u1 = uniform01(n, seed)
    u2 = uniform01(n, seed ^ 0x5A5A5A5A)
    x = sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    r = splitmix64(seed ^ 0xA5A5A5A5, n)
    sel = (r % np.uint64(100)).astype(np.int64)
    sign = ((r >> np.uint64(8)) & np.uint64(1)).astype(bool)
    bits = x.view(np.uint64).copy()
    zero = sel == 0
    bits[zero] = np.where(sign[zero], np.uint64(0x8000000000000000), np.uint64(0))
    sub = sel == 1
    mant = (r[sub] >> np.uint64(9)) & np.uint64((1 << 52) - 1)
    mant[mant == 0] = 1
    bits[sub] = mant | np.where(sign[sub], np.uint64(0x8000000000000000), np.uint64(0))
    big = sel == 2
    bits[big] = np.where(sign[big], np.float64(-1e300), np.float64(1e300)).view(np.uint64)
    return bits.view(np.float64)


def floats_finite(n: int, seed: int) -> np.ndarray:
    """Random finite float32 bit patterns (exponent != 255)."""
    b = splitmix64(seed, n).astype(np.uint32)
    exp = (b >> np.uint32(23)) & np.uint32(0xFF)
    b[exp == 0xFF] &= np.uint32(0xFF7FFFFF)
    return b.view(np.float32)


def segment_lengths(nseg: int, seed: int, lo: int = 16, hi: int = 1 << 20) -> np.ndarray:
    """floor(exp(U(ln lo, ln hi))) segment lengths (log-uniform)."""
    u = uniform01(nseg, seed)
    return np.floor(np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))).astype(np.int64)


# ---- the same generators on a torch device (GPU parity tests at full size) ----
# torch has no uint64 arithmetic: int64 holds the same bits, multiplication
# wraps modulo 2^64, and a logical right shift is the arithmetic shift masked
# to the low 64 - k bits.  The integer outputs equal splitmix64's bit for bit.

def _as_i64(x: int) -> int:
    x &= 0xFFFFFFFFFFFFFFFF
    return x - (1 << 64) if x >= 1 << 63 else x


def splitmix64_torch(seed: int, n: int, device="cuda", chunk: int = 1 << 26):
    """int64 tensor holding the bits of splitmix64(seed) outputs [0, n)."""
    import torch

    def srl(z, k):
        return (z >> k) & ((1 << (64 - k)) - 1)

    out = torch.empty(n, dtype=torch.int64, device=device)
    golden, m1, m2 = _as_i64(int(_GOLDEN)), _as_i64(int(_M1)), _as_i64(int(_M2))
    s = _as_i64(seed)
    for lo in range(0, n, chunk):
        hi = min(n, lo + chunk)
        z = torch.arange(lo + 1, hi + 1, dtype=torch.int64, device=device)
        z.mul_(golden).add_(s)
        z ^= srl(z, 30)
        z.mul_(m1)
        z ^= srl(z, 27)
        z.mul_(m2)
        z ^= srl(z, 31)
        out[lo:hi] = z
    return out


def keys_u32_torch(n: int, seed: int, dist: str = "uniform", device="cuda"):
    """keys_u32 as an int32 tensor of the same bits."""
    import torch
    if dist == "sorted":
        return torch.arange(n, dtype=torch.int64, device=device).to(torch.int32)
    if dist == "constant":
        return torch.zeros(n, dtype=torch.int32, device=device)
    low = splitmix64_torch(seed, n, device) & 0xFFFFFFFF
    if dist == "low8":
        c = int(splitmix64(seed ^ 0xC0FFEE, 1)[0]) & 0xFFFFFF00
        low = (low & 0xFF) | c
    elif dist != "uniform":
        raise ValueError(f"unknown distribution {dist}")
    return (low - ((low >> 31) << 32)).to(torch.int32)   # the low 32 bits as a signed int32
