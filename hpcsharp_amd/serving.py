"""Serving loop: trace generation, the run through libhpcsort.so's
hpcs_serve_run, and the latency / throughput summary.

The workload: Poisson arrivals at `rate` requests/s; each request sorts or
takes the top k of a payload of n keys (u32, f32 or u64) generated on the
device from a per-request seed.  The engine itself (admission against the HBM
budget, worker slots with one stream each, completion tracking) is C++ in
hpcsharp_amd/csrc/hpcs_serve.hip; this module builds the request table and
reads the results.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from hpcsharp_amd import _native as N
from hpcsharp_amd import datagen

REQUEST = np.dtype([("op", np.int32), ("dtype", np.int32), ("n", np.int64), ("k", np.int64),
                    ("seed", np.uint64), ("arrival_s", np.float64)])
RESULT = np.dtype([("admit_s", np.float64), ("enqueued_s", np.float64), ("done_s", np.float64), ("checksum", np.uint64),
                   ("status", np.int32), ("slot", np.int32)])


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("keys", ctypes.c_int64), ("peak_bytes", ctypes.c_int64),
                ("request_budget_bytes", ctypes.c_int64), ("scratch_bytes_per_slot", ctypes.c_int64),
                ("slots", ctypes.c_int32), ("pad", ctypes.c_int32), ("batches", ctypes.c_int64),
                ("coalesced_requests", ctypes.c_int64)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int32), ("mode", ctypes.c_int32), ("hbm_budget", ctypes.c_int64),
                ("request_timeout_s", ctypes.c_double), ("coalesce_max_keys", ctypes.c_int64),
                ("coalesce_max_requests", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("coalesce_max_batch_keys", ctypes.c_int64)]


DTYPES = {N.HPCS_DTYPE_U32: np.uint32, N.HPCS_DTYPE_U64: np.uint64,
          N.HPCS_DTYPE_F32: np.float32, N.HPCS_DTYPE_F64: np.float64}
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def make_trace(nreq: int = 20000, rate: float = 2000.0, seed: int = 0x5EED0009, sort_frac: float = 0.7,
               n_lo: int = 1 << 10, n_hi: int = 1 << 26, k_lo: int = 10, k_hi: int = 10 ** 4,
               dtypes=(N.HPCS_DTYPE_U32, N.HPCS_DTYPE_F32, N.HPCS_DTYPE_U64)) -> np.ndarray:
    """The serving workload: exponential gaps of mean 1/rate; op sort with
    probability sort_frac, else top-k with k = floor(exp(U(ln k_lo, ln k_hi)))
    (at most n); n = floor(exp(U(ln n_lo, ln n_hi))); dtype uniform over
    `dtypes`.  Every draw comes from splitmix64 streams of `seed`."""
    u = [datagen.uniform01(nreq, seed ^ (0x1111 * (j + 1))) for j in range(5)]
    t = np.empty(nreq, dtype=REQUEST)
    t["arrival_s"] = np.cumsum(-np.log(u[0]) / rate)
    t["op"] = np.where(u[1] < sort_frac, N.HPCS_OP_SORT, N.HPCS_OP_TOPK)
    t["n"] = np.floor(np.exp(np.log(n_lo) + u[2] * (np.log(n_hi) - np.log(n_lo)))).astype(np.int64)
    k = np.floor(np.exp(np.log(k_lo) + u[3] * (np.log(k_hi) - np.log(k_lo)))).astype(np.int64)
    t["k"] = np.where(t["op"] == N.HPCS_OP_TOPK, np.minimum(k, t["n"]), 0)
    t["dtype"] = np.asarray(dtypes, dtype=np.int32)[np.minimum((u[4] * len(dtypes)).astype(np.int64), len(dtypes) - 1)]
    t["seed"] = datagen.splitmix64(seed ^ 0x5E5E5E5E, nreq)
    return t


def run(trace: np.ndarray, slots: int = 4, hbm_budget: int = 200 * 10 ** 9,
        mode: int = N.HPCS_SERVE_REALTIME, device: int = 0, timeout_s: float = 0.0,
        coalesce_max_keys: int = 0, coalesce_max_requests: int = 0, coalesce_max_batch_keys: int = 0):
    """Replay the trace on one GPU; returns (results, stats).  timeout_s: the
    longest wait for one request's completion (0: 5 s + 20 ns per key);
    coalesce_max_keys > 0: queued u32 / f32 sorts of at most that many keys
    run together as one segmented sort (coalesce_max_requests per batch, 0:
    256; coalesce_max_batch_keys keys per batch, 0: 2^22)."""
    trace = np.ascontiguousarray(trace, dtype=REQUEST)
    if not os.environ.get("HPCS_TUNE_FILE") and os.path.exists(N.TUNE_PATH):
        os.environ["HPCS_TUNE_FILE"] = N.TUNE_PATH   # the slots' contexts load the measured table
    res = np.zeros(trace.size, dtype=RESULT)
    st = ServeStats()
    cfg = ServeConfig(slots=slots, mode=mode, hbm_budget=int(hbm_budget), request_timeout_s=float(timeout_s),
                      coalesce_max_keys=int(coalesce_max_keys), coalesce_max_requests=int(coalesce_max_requests),
                      coalesce_max_batch_keys=int(coalesce_max_batch_keys))
    rc = N.lib().hpcs_serve_run_config(device, ctypes.byref(cfg), trace.ctypes.data_as(ctypes.c_void_p), trace.size,
                                       res.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st))
    if rc != N.HPCS_OK:
        raise N.HpcsError(rc, N.lib().hpcs_serve_last_error().decode())
    return res, st


def payload(req) -> np.ndarray:
    """Host copy of a request's payload (the device generator's output)."""
    x = datagen.splitmix64(int(req["seed"]), int(req["n"]))
    dt = int(req["dtype"])
    if dt in (N.HPCS_DTYPE_U32, N.HPCS_DTYPE_F32):
        b = x.astype(np.uint32)
        if dt == N.HPCS_DTYPE_F32:
            inf = ((b >> np.uint32(23)) & np.uint32(0xFF)) == np.uint32(0xFF)
            b[inf] &= np.uint32(0xFF7FFFFF)
            return b.view(np.float32)
        return b
    if dt == N.HPCS_DTYPE_F64:
        inf = ((x >> np.uint64(52)) & np.uint64(0x7FF)) == np.uint64(0x7FF)
        x[inf] &= ~np.uint64(1 << 52)
        return x.view(np.float64)
    return x


def checksum(keys: np.ndarray) -> int:
    """sum_i mix64(bits_i + (i + 1) * golden) mod 2^64 (the engine's result
    checksum)."""
    bits = keys.view(np.uint32 if keys.itemsize == 4 else np.uint64).astype(np.uint64)
    with np.errstate(over="ignore"):
        z = bits + (np.arange(1, bits.size + 1, dtype=np.uint64) * _GOLDEN)
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
        return int(np.sum(z, dtype=np.uint64))


def summarize(trace: np.ndarray, res: np.ndarray, stats: ServeStats | None = None) -> dict:
    """Latency (arrival to completion) and admission-wait percentiles in ms,
    throughput in Gkeys/s over the run's wall time."""
    ok = res["status"] == N.HPCS_OK
    lat = (res["done_s"] - trace["arrival_s"])[ok] * 1e3
    wait = (res["admit_s"] - trace["arrival_s"])[ok] * 1e3
    enq = (res["enqueued_s"] - res["admit_s"])[ok] * 1e3
    keys = int(trace["n"][ok].sum())
    wall = float(stats.wall_s) if stats is not None else float(res["done_s"].max())
    span = float(trace["arrival_s"][-1]) if trace.size else 0.0
    out = {
        "requests": int(trace.size), "completed": int(ok.sum()), "rejected": int((~ok).sum()),
        "p50_ms": float(np.percentile(lat, 50)) if lat.size else None,
        "p99_ms": float(np.percentile(lat, 99)) if lat.size else None,
        "mean_ms": float(lat.mean()) if lat.size else None,
        "admit_wait_p50_ms": float(np.percentile(wait, 50)) if wait.size else None,
        "admit_wait_p99_ms": float(np.percentile(wait, 99)) if wait.size else None,
        "enqueue_p50_ms": float(np.percentile(enq, 50)) if enq.size else None,
        "gkeys_per_s": keys / wall / 1e9 if wall > 0 else None,
        "offered_gkeys_per_s": keys / span / 1e9 if span > 0 else None,
        "wall_s": wall,
    }
    for name, op in (("sort", N.HPCS_OP_SORT), ("topk", N.HPCS_OP_TOPK)):
        sel = (trace["op"] == op)[ok]
        if sel.any():
            out[f"{name}_p50_ms"] = float(np.percentile(lat[sel], 50))
            out[f"{name}_p99_ms"] = float(np.percentile(lat[sel], 99))
    if stats is not None:
        out.update(slots=int(stats.slots), peak_request_bytes=int(stats.peak_bytes),
                   request_budget_bytes=int(stats.request_budget_bytes),
                   scratch_bytes_per_slot=int(stats.scratch_bytes_per_slot),
                   batches=int(stats.batches), coalesced_requests=int(stats.coalesced_requests))
    return out
