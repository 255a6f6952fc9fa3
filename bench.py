#!/usr/bin/env python3
"""Benchmark: stable LSD radix sort of 2^28 uint32 keys on one MI355X.

One step = one full device-resident sort (upsweep, plan, then per 8-bit
digit pass: tile counts, scan, scatter) of a distinct copy of the same
2^28-key array (uniform 32-bit keys from splitmix64, seed 0x5EED0002),
already resident in HBM.  W untimed
warm-up sorts, then exactly K timed sorts bracketed by device
synchronisation; every timed output is checked against an independent sort
(torch.sort) before anything is printed.

Prints one JSON line: Gkeys/s (the K timed sorts, no events among them), the
pass scatter kernel's roofline (algorithmic bytes = 8 B per key per pass,
divided by its launch duration from HIP events on the launch stream, recorded
over a second region of K sorts) and the
CPU baseline (the C restatement of ParallelAlgorithm.SortRadixPar, timed on
this host's cores).  Per-distribution figures go to stderr and to the JSON
line's "extra" field.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
N_DEFAULT = 1 << 28
SEED = 0x5EED0002


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_sorts(torch, N, src, copies: int):
    bufs = [src.clone() for _ in range(copies)]
    torch.cuda.synchronize()
    return bufs


def run_sort(N, lib, ctx, t, n, stream):
    N.check(lib.hpcs_dev_sort_radix_u32(ctx, ctypes.c_void_p(t.data_ptr()), n, ctypes.c_void_p(stream)))


def measure(torch, N, host_keys: np.ndarray, steps: int, warmup: int, profile: bool):
    """Seconds per sort over `steps` timed sorts of distinct device-resident
    copies (no events inside that region); with `profile`, a second region of
    `steps` sorts brackets every kernel with HIP events on the launch stream
    (the per-kernel times; the events themselves add about 0.1 ms per sort,
    so they stay out of the first region).  Returns (s per sort, per-kernel
    stats, s per sort of the evented region, output correct)."""
    lib, ctx = N.lib(), N.context(0)
    n = host_keys.size
    src = torch.from_numpy(host_keys.view(np.int32)).cuda()
    regions = 2 if profile else 1
    bufs = gpu_sorts(torch, N, src, warmup + regions * steps)
    stream = torch.cuda.current_stream().cuda_stream
    for i in range(warmup):
        run_sort(N, lib, ctx, bufs[i], n, stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        run_sort(N, lib, ctx, bufs[warmup + i], n, stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stats, t_ev = {}, None
    if profile:
        N.profile_enable(True)
        t2 = time.perf_counter()
        for i in range(steps):
            run_sort(N, lib, ctx, bufs[warmup + steps + i], n, stream)
        torch.cuda.synchronize()
        t_ev = (time.perf_counter() - t2) / steps
        stats = N.profile_read()
        N.profile_enable(False)
    # correctness before reporting: every timed output equals an independent
    # library sort
    ref = torch.sort(src.to(torch.int64) & 0xFFFFFFFF)[0]
    ok = all(bool(torch.equal(b.to(torch.int64) & 0xFFFFFFFF, ref)) for b in bufs[warmup:])
    del bufs, ref, src
    torch.cuda.empty_cache()
    return (t1 - t0) / steps, stats, t_ev, ok


def batched_vs_plain(torch, N, datagen, nseg: int, seed: int, reps: int):
    """4096 uint32 segments with lengths log-uniform in [16, 2^20] (uniform
    keys): one batched call (captured once in a HIP graph, replayed) against
    a loop of plain per-array calls on the same buffers.  Outputs must be
    identical.  Returns a dict for the JSON line."""
    lens = datagen.segment_lengths(nseg, seed, 16, 1 << 20)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(begin[-1])
    keys = datagen.keys_u32(total, seed + 1, "uniform")
    lib, ctx = N.lib(), N.context(0)
    src = torch.from_numpy(keys.view(np.int32)).cuda()
    d_begin = torch.from_numpy(begin).cuda()
    s = torch.cuda.Stream()
    work = src.clone()
    P = ctypes.c_void_p

    def batched():
        N.check(lib.hpcs_dev_sort_segments_u32(ctx, P(work.data_ptr()), total, P(d_begin.data_ptr()), nseg,
                                               P(s.cuda_stream)))

    with torch.cuda.stream(s):
        batched()                                     # warm-up: sizes the scratch
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        batched()
    # the replay between HIP events on the stream it is launched on
    tb = event_ms(torch, lambda: work.copy_(src), g.replay, reps) * 1e-3
    out_b = work.clone()
    # plain loop: one device-resident sort call per segment
    plain = src.clone()
    base = plain.data_ptr()
    t_p = []
    for _ in range(max(1, reps // 3)):
        plain.copy_(src)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(nseg):
            n_i = int(lens[i])
            if n_i > 1:
                N.check(lib.hpcs_dev_sort_radix_u32(ctx, P(base + 4 * int(begin[i])), n_i, P(s.cuda_stream)))
        torch.cuda.synchronize()
        t_p.append(time.perf_counter() - t0)
    identical = bool(torch.equal(out_b, plain))
    tp = float(np.median(t_p))
    del work, out_b, plain, src, g
    torch.cuda.empty_cache()
    return {"workload": f"{nseg} uint32 segments, lengths log-uniform in [16, 2^20], uniform keys, seed {seed:#x}",
            "keys": total, "batched_ms": round(tb * 1e3, 3), "plain_loop_ms": round(tp * 1e3, 3),
            "batched_gkeys": round(total / tb / 1e9, 3), "plain_loop_gkeys": round(total / tp / 1e9, 3),
            "speedup": round(tp / tb, 2), "identical_outputs": identical,
            "batched": "one hipGraph replay of the batched call (its fixed launch sequence), HIP events around "
                       "the replay; the plain loop timed on the host"}


def timed_calls(torch, fn, reps: int) -> float:
    """Seconds per call of fn() (device work on the current stream)."""
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def msd_split_figures(torch, N, reps: int):
    """The opt-in MSD split (HPCS_OPT_PASS_ALGO 4) beside the default (tuned)
    sort on uniform keys, device-resident, event medians; outputs compared."""
    lib = N.lib()
    P = ctypes.c_void_p
    cd = N.new_context(0)
    cm = N.new_context(0)
    out = {}
    try:
        for ctx in (cd, cm):
            N.load_tuning(N.TUNE_PATH, ctx)
        N.set_option(N.HPCS_OPT_PASS_ALGO, 4, cm)
        s = torch.cuda.Stream()
        for name, kb, fn in (("u32", 4, lib.hpcs_dev_sort_radix_u32), ("u64", 8, lib.hpcs_dev_sort_radix_u64)):
            for lg in (16, 20):
                n = 1 << lg
                g = torch.Generator(device="cuda").manual_seed(lg + kb)
                src = torch.randint(-2 ** 31, 2 ** 31 - 1, (n * kb // 4,), dtype=torch.int32, device="cuda",
                                    generator=g)
                res, outs = [], []
                with torch.cuda.stream(s):
                    for ctx in (cd, cm):
                        work = src.clone()
                        call = lambda: N.check(fn(ctx, P(work.data_ptr()), n, P(s.cuda_stream)))
                        res.append(event_ms(torch, lambda: work.copy_(src), call, reps))
                        outs.append(work)
                s.synchronize()
                out[f"{name}_2p{lg}"] = {"default_us": round(res[0] * 1e3, 1), "msd_us": round(res[1] * 1e3, 1),
                                         "identical": bool(torch.equal(outs[0], outs[1]))}
    finally:
        N.destroy_context(cd)
        N.destroy_context(cm)
    out["note"] = "uniform keys; clustered keys make the MSD split several times slower, so it is opt-in"
    return out


def topk_figures(torch, N, datagen, reps: int):
    """Radix top-k, k = 10^4 of 2^26 keys (the serving workload's largest
    top-k requests), per key type; the u32 result is checked on the host."""
    lib, ctx = N.lib(), N.context(0)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n, k = 1 << 26, 10 ** 4
    out = {}
    for name, fn, gen in (("u32", "hpcs_dev_topk_u32", lambda: datagen.keys_u32(n, 0x5EED0009)),
                          ("f32", "hpcs_dev_topk_f32", lambda: datagen.floats_finite(n, 0x5EED0009).view(np.uint32)),
                          ("u64", "hpcs_dev_topk_u64", lambda: datagen.keys_u64(n, 0x5EED0009))):
        h = gen()
        x = torch.from_numpy(h.view(np.int32 if h.itemsize == 4 else np.int64)).cuda()
        o = torch.empty(k, dtype=x.dtype, device="cuda")
        f = getattr(lib, fn)
        call = lambda: N.check(f(ctx, ctypes.c_void_p(x.data_ptr()), n, k, ctypes.c_void_p(o.data_ptr()), st))
        t = timed_calls(torch, call, reps)
        if name == "u32":
            want = np.sort(np.partition(h, n - k)[n - k:])[::-1]
            if not np.array_equal(o.cpu().numpy().view(np.uint32), want):
                print("bench: top-k output differs from the host selection", file=sys.stderr)
                sys.exit(3)
        out[name] = {"ms": round(t * 1e3, 4), "gkeys_per_s": round(n / t / 1e9, 2),
                     "read_GBs": round(n * h.itemsize / t / 1e9, 1)}
        del x, o
    return {"n": n, "k": k, "by_type": out}


def event_ms(torch, prep, fn, reps: int) -> float:
    """Median milliseconds of fn() between HIP events on the current stream,
    prep() run before each call outside the timed interval."""
    st = torch.cuda.current_stream()
    prep()
    fn()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        prep()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        fn()
        b.record(st)
        b.synchronize()
        ms.append(a.elapsed_time(b))
    return float(np.median(ms))


def histogram_figures(torch, N, datagen, keys_host: np.ndarray, reps: int):
    """HistogramByteComponents (device-resident): on the headline array
    (GB/s against the one-read roofline, counts checked with torch.bincount)
    and on the 2^20-key array of the reference's CPU-runnable workload (seed
    0x5EED0001), together with that array's sort latency."""
    lib, ctx = N.lib(), N.context(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = ctypes.c_void_p
    out = {}
    for tag, h in (("n_2p28", keys_host), ("n_2p20", datagen.keys_u32(1 << 20, 0x5EED0001))):
        n = h.size
        x = torch.from_numpy(h.view(np.int32)).cuda()
        counts = torch.empty(4 * 256, dtype=torch.int32, device="cuda")
        call = lambda: N.check(lib.hpcs_dev_histogram_byte_components_u32(ctx, P(x.data_ptr()), n,
                                                                          P(counts.data_ptr()), sp))
        ms = event_ms(torch, lambda: None, call, reps)
        xl = x.to(torch.int64) & 0xFFFFFFFF
        want = torch.stack([torch.bincount((xl >> (8 * d)) & 255, minlength=256) for d in range(4)]).reshape(-1)
        if not torch.equal(counts.to(torch.int64), want):
            print("bench: histogram differs from torch.bincount", file=sys.stderr)
            sys.exit(3)
        r = {"n": n, "histogram_us": round(ms * 1e3, 2), "histogram_gkeys_per_s": round(n / ms / 1e6, 2),
             "histogram_GBs": round(4 * n / ms / 1e6, 1)}
        if tag == "n_2p28":
            r["histogram_roofline_frac"] = round(4 * n / ms / 1e6 / HBM_PEAK_GBS, 4)
        else:
            work = x.clone()
            sort = lambda: N.check(lib.hpcs_dev_sort_radix_u32(ctx, P(work.data_ptr()), n, sp))
            sms = event_ms(torch, lambda: work.copy_(x), sort, reps)
            ref, _ = torch.sort(xl)
            if not torch.equal(work.to(torch.int64) & 0xFFFFFFFF, ref):
                print("bench: 2^20 sort differs from torch.sort", file=sys.stderr)
                sys.exit(3)
            r["sort_us"] = round(sms * 1e3, 2)
            r["sort_gkeys_per_s"] = round(n / sms / 1e6, 2)
            # the same sort through a context's graph cache (HPCS_OPT_GRAPHS):
            # repeated calls launch one captured graph of the sequence
            gctx = N.new_context(0)
            N.load_tuning(N.TUNE_PATH, gctx)
            N.set_option(N.HPCS_OPT_GRAPHS, 1, gctx)
            gs = torch.cuda.Stream()
            with torch.cuda.stream(gs):
                gsort = lambda: N.check(lib.hpcs_dev_sort_radix_u32(gctx, P(work.data_ptr()), n, P(gs.cuda_stream)))
                work.copy_(x)
                gsort()                                   # first call: direct
                gms = event_ms(torch, lambda: work.copy_(x), gsort, reps)
            gs.synchronize()
            if not torch.equal(work.to(torch.int64) & 0xFFFFFFFF, ref):
                print("bench: graph-replayed 2^20 sort differs from torch.sort", file=sys.stderr)
                sys.exit(3)
            _, caps, replays = N.graph_cache_stats(gctx)
            N.destroy_context(gctx)
            r["sort_graph_us"] = round(gms * 1e3, 2)
            r["sort_graph_replays"] = replays
            del work
        out[tag] = r
        del x, counts, xl, want
    return out


def histogram_cases(torch, N, datagen, reps: int):
    """The byte histogram at 2^28 on ascending and constant u32 keys and on
    uniform u64 keys (bench.py --only histogram), counts checked with
    torch.bincount."""
    lib, ctx = N.lib(), N.context(0)
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = ctypes.c_void_p
    n = 1 << 28
    out = {}
    for tag, gen, fn, kb in (("u32_sorted", lambda: datagen.keys_u32(n, SEED, "sorted"),
                              "hpcs_dev_histogram_byte_components_u32", 4),
                             ("u32_constant", lambda: datagen.keys_u32(n, SEED, "constant"),
                              "hpcs_dev_histogram_byte_components_u32", 4),
                             ("u64_uniform", lambda: datagen.keys_u64(n, SEED),
                              "hpcs_dev_histogram_byte_components_u64", 8)):
        h = gen()
        x = torch.from_numpy(h.view(np.int32 if kb == 4 else np.int64)).cuda()
        del h
        counts = torch.empty(kb * 256, dtype=torch.int32, device="cuda")
        f = getattr(lib, fn)
        ms = event_ms(torch, lambda: None,
                      lambda: N.check(f(ctx, P(x.data_ptr()), n, P(counts.data_ptr()), sp)), reps)
        xl = x.to(torch.int64)
        want = torch.stack([torch.bincount((xl >> (8 * d)) & 255, minlength=256) for d in range(kb)]).reshape(-1)
        if not torch.equal(counts.to(torch.int64), want):
            print(f"bench: {tag} histogram differs from torch.bincount", file=sys.stderr)
            sys.exit(3)
        out[tag] = {"n": n, "histogram_us": round(ms * 1e3, 2), "histogram_GBs": round(kb * n / ms / 1e6, 1),
                    "histogram_roofline_frac": round(kb * n / ms / 1e6 / HBM_PEAK_GBS, 4)}
        del x, counts, xl, want
    return out


def cfg3_figures(torch, N, datagen, reps: int):
    """Floating and record keys at 2^27: double[] with +-0.0, subnormals and
    +-1e300 (plain sort), and {u64 key, u32 payload} records through the
    key-offset record sort (stable; 8 passes over the keys)."""
    lib, ctx = N.lib(), N.context(0)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = 1 << 27
    out = {}
    d = datagen.doubles_special(n, 0x5EED0006)
    src = torch.from_numpy(d.view(np.int64)).cuda()
    work = src.clone()

    def f64():
        work.copy_(src)
        N.check(lib.hpcs_dev_sort_radix_f64(ctx, ctypes.c_void_p(work.data_ptr()), n, st))
    t_copy = timed_calls(torch, lambda: work.copy_(src), reps)
    t = timed_calls(torch, f64, reps) - t_copy
    # bit-exact against an independent sort: torch.sort of the radix keys
    # (negative: all bits flipped; others: sign bit flipped) as order-
    # preserving int64, mapped back to double bits; -0.0 sorts before +0.0
    sign = -(1 << 63)
    neg = src < 0
    key = torch.where(neg, ~src, src ^ sign) ^ sign          # radix key, in signed order
    want_key = torch.sort(key)[0] ^ sign                      # radix keys ascending (bits)
    want = torch.where(want_key < 0, want_key ^ sign, ~want_key)   # inverse transform
    ok = bool(torch.equal(work, want))
    del neg, key, want_key, want
    out["f64_sort"] = {"ms": round(t * 1e3, 3), "gkeys_per_s": round(n / t / 1e9, 2), "bit_exact": ok}
    if not ok:
        print("bench: f64 output differs from the independent sort", file=sys.stderr)
        sys.exit(3)
    del src, work
    rec = np.zeros(n, dtype=np.dtype({"names": ["key", "payload"], "formats": [np.uint64, np.uint32],
                                      "offsets": [0, 8], "itemsize": 16}))
    rec["key"] = datagen.keys_u64(n, 0x5EED0007)
    rec["payload"] = np.arange(n, dtype=np.uint32)
    rsrc = torch.from_numpy(rec.view(np.uint8).reshape(-1)).cuda()
    rwork = rsrc.clone()

    def recs():
        rwork.copy_(rsrc)
        N.check(lib.hpcs_dev_sort_radix_records(ctx, ctypes.c_void_p(rwork.data_ptr()), n, 16, 0, 64,
                                                N.HPCS_KEY_UINT, st))
    t_copy = timed_calls(torch, lambda: rwork.copy_(rsrc), reps)
    t = timed_calls(torch, recs, reps) - t_copy
    r = rwork.cpu().numpy().view(rec.dtype)
    ties = r["key"][1:] == r["key"][:-1]
    ok = bool(np.all(r["key"][1:] >= r["key"][:-1])) and bool(
        np.array_equal(np.sort(r["payload"]), rec["payload"])) and bool(
        np.array_equal(rec["key"][r["payload"]], r["key"])) and bool(
        np.all(r["payload"][1:][ties] > r["payload"][:-1][ties]))
    out["records_u64_key_u32_payload"] = {"ms": round(t * 1e3, 3), "gkeys_per_s": round(n / t / 1e9, 2),
                                          "sorted_stable_permutation": ok}
    if not ok:
        print("bench: record sort is not a stable sort of the input", file=sys.stderr)
        sys.exit(3)
    return {"n": n, **out}


def cfg3_cpu_baseline(datagen, lg: int = 24):
    """The floating and record workloads on the host: single-threaded C
    restatements of SortRadixMsd(double[]) and SortRadixLsd2 over records
    (oracle/), on a 2^lg sample of the same generators; best of 2."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    n = 1 << lg
    d = datagen.doubles_special(n, 0x5EED0006)
    rec = np.zeros(n, dtype=np.dtype({"names": ["key", "payload"], "formats": [np.uint64, np.uint32],
                                      "offsets": [0, 8], "itemsize": 16}))
    rec["key"] = datagen.keys_u64(n, 0x5EED0007)
    rec["payload"] = np.arange(n, dtype=np.uint32)
    out = {"sample": f"2^{lg} elements of the same generators, 1 thread, best of 2, on {cpu_model()}",
           "cores": 1, "kind": "port"}
    for name, src, fn in (("f64_sort", d, O.sort_radix_float),
                          ("records_u64_key_u32_payload", rec, lambda r: O.sort_radix_records(r, 0, 64))):
        best = float("inf")
        for _ in range(2):
            a = src.copy()
            t0 = time.perf_counter()
            fn(a)
            best = min(best, time.perf_counter() - t0)
        out[name] = {"ms": round(best * 1e3, 1), "gkeys_per_s": round(n / best / 1e9, 4)}
    return out


def host_array_figures(N, keys: np.ndarray, reps: int):
    """The host-array drop-in call (hpcs_sort_radix_u32 on pageable host
    memory): PCIe copy in, sort, copy out; never the headline value."""
    lib, ctx = N.lib(), N.context(0)
    times = []
    for r in range(reps + 1):
        a = keys.copy()
        t0 = time.perf_counter()
        N.check(lib.hpcs_sort_radix_u32(ctx, a.ctypes.data_as(ctypes.c_void_p), a.size))
        if r:
            times.append(time.perf_counter() - t0)
    ok = bool((a[1:] >= a[:-1]).all())
    t = float(np.median(times))
    return {"ms": round(t * 1e3, 2), "gkeys_per_s": round(keys.size / t / 1e9, 3), "sorted": ok,
            "note": "includes host-to-device and device-to-host copies of the whole array"}


COALESCE_MAX_KEYS = 1 << 16   # serving: u32 / f32 sorts up to this size may share a segmented sort


def serving_figures(nreq: int, slots: int, sweep: bool = True):
    """The serving workload: Poisson arrivals (2000/s), 70% sort / 30% top-k,
    n log-uniform in [2^10, 2^26), u32/f32/u64, 200 GB budget; once as is,
    once with small sorts coalesced, then (sweep) the coalescing engine at
    1x, 2x, 4x, 8x the arrival rate.  Sizes span 2^16, so p99 is far above
    p50 even on an idle GPU: saturation is judged against the 1x run, as the
    highest rate whose p99 stays within 10x the 1x p99 while the achieved
    throughput keeps within 5% of the offered.  Every request's checksum must
    be the same in every run (the same payloads and results)."""
    from hpcsharp_amd import serving as S

    def fig(t, **kw):
        res, st = S.run(t, slots=slots, hbm_budget=200 * 10 ** 9, **kw)
        s = S.summarize(t, res, st)
        return res, {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()}

    t = S.make_trace(nreq=nreq)
    res0, out = fig(t)
    res1, out["coalesced"] = fig(t, coalesce_max_keys=COALESCE_MAX_KEYS)
    out["coalesced"]["same_checksums"] = bool(np.array_equal(res0["checksum"], res1["checksum"]))
    if sweep:
        rows, best, base_p99 = [], None, None
        for mult in (1, 2, 4, 8):
            tm = S.make_trace(nreq=nreq, rate=2000.0 * mult)
            resm, f = fig(tm, coalesce_max_keys=COALESCE_MAX_KEYS)
            row = {"rate_per_s": 2000 * mult, "p50_ms": f["p50_ms"], "p99_ms": f["p99_ms"],
                   "gkeys_per_s": f["gkeys_per_s"], "offered_gkeys_per_s": f["offered_gkeys_per_s"],
                   "coalesced_requests": f["coalesced_requests"], "batches": f["batches"],
                   "completed": f["completed"], "same_checksums": bool(np.array_equal(resm["checksum"], res0["checksum"]))}
            rows.append(row)
            base_p99 = base_p99 or f["p99_ms"]
            if f["p99_ms"] <= 10 * base_p99 and f["gkeys_per_s"] >= 0.95 * f["offered_gkeys_per_s"]:
                best = row
        out["saturation_sweep"] = {
            "rows": rows,
            "highest_sustained_gkeys_per_s": best["gkeys_per_s"] if best else None,
            "sustained_rate_per_s": best["rate_per_s"] if best else None,
            "peak_gkeys_per_s": max(r["gkeys_per_s"] for r in rows),
            "note": "coalescing engine at 1x-8x of 2000 requests/s, the same seeds at every rate (checksums must "
                    "match); sustained: the highest rate whose p99 stays within 10x the 1x p99 with the achieved "
                    "throughput within 5% of the offered; peak: the most Gkeys/s achieved at any rate"}
    return out


def cpu_model() -> str:
    """Model name of the host CPU, from /proc/cpuinfo ("unknown CPU" elsewhere)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(host_keys: np.ndarray, reps: int):
    """ParallelAlgorithm.SortRadixPar restated in C (oracle/), OpenMP threads."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    lib = O.lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    a = host_keys.copy()
    O.sort_radix_par(a)           # warm-up
    times = []
    for _ in range(reps):
        a = host_keys.copy()
        t0 = time.perf_counter()
        O.sort_radix_par(a)
        times.append(time.perf_counter() - t0)
    assert bool((a[1:] >= a[:-1]).all())
    del lib
    t = float(np.median(times))
    return host_keys.size / t / 1e9, threads, times


def segments_only(torch, N, datagen, nseg: int, seed: int, reps: int):
    """The batched call alone (4096 segments, lengths log-uniform in [16,
    2^20]), event medians over `reps` calls on freshly copied keys; the output
    is checked against torch.sort of (segment, key) pairs."""
    lens = datagen.segment_lengths(nseg, seed, 16, 1 << 20)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(begin[-1])
    src = datagen.keys_u32_torch(total, seed + 1)
    d_begin = torch.from_numpy(begin).cuda()
    work = src.clone()
    lib, ctx = N.lib(), N.context(0)
    P = ctypes.c_void_p
    st = P(torch.cuda.current_stream().cuda_stream)
    call = lambda: N.check(lib.hpcs_dev_sort_segments_u32(ctx, P(work.data_ptr()), total, P(d_begin.data_ptr()),
                                                          nseg, st))
    ms = event_ms(torch, lambda: work.copy_(src), call, reps)
    seg = torch.repeat_interleave(torch.arange(nseg, dtype=torch.int64, device="cuda"), torch.from_numpy(lens).cuda())
    want = torch.sort((seg << 32) | (src.to(torch.int64) & 0xFFFFFFFF))[0] & 0xFFFFFFFF
    ok = bool(torch.equal(work.to(torch.int64) & 0xFFFFFFFF, want))
    if not ok:
        print("bench: batched segments differ from torch.sort", file=sys.stderr)
        sys.exit(3)
    return {"keys": total, "segments": nseg, "ms": round(ms, 4), "gkeys_per_s": round(total / ms / 1e6, 2)}


def only_section(torch, N, datagen, args):
    """One workload alone (bench.py --only X), for profiling runs."""
    if args.only == "histogram":
        r = histogram_figures(torch, N, datagen, datagen.keys_u32(args.n, SEED, "uniform"), 10)
        r.update(histogram_cases(torch, N, datagen, 10))
        return r
    if args.only == "topk":
        return topk_figures(torch, N, datagen, 5)
    if args.only == "segments":
        return segments_only(torch, N, datagen, 4096, 0x5EED0008, 5)
    if args.only == "serving":
        return serving_figures(args.serving_requests, 4)
    if args.only == "cfg3":
        return cfg3_figures(torch, N, datagen, 3)
    return msd_split_figures(torch, N, 10)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dists", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-topk", action="store_true")
    ap.add_argument("--no-host", action="store_true")
    ap.add_argument("--no-cfg3", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-hist", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=20000)
    ap.add_argument("--only", choices=("sort", "histogram", "topk", "segments", "serving", "cfg3", "msd"),
                    help="run one workload alone (for profiling runs): 'sort' is the headline sort's line "
                         "without the other workloads; the others print that workload's figures as JSON")
    args = ap.parse_args()
    if args.only == "sort":
        args.no_cpu = args.no_dists = args.no_batched = args.no_topk = args.no_host = True
        args.no_cfg3 = args.no_serving = args.no_hist = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (use --gpus 1).", file=sys.stderr)
        sys.exit(2)

    import torch
    from hpcsharp_amd import _native as N
    from hpcsharp_amd import datagen

    if not torch.cuda.is_available():
        print("no HIP device", file=sys.stderr)
        sys.exit(1)
    n = args.n
    if args.only not in (None, "sort"):
        print(json.dumps({"workload": args.only, **only_section(torch, N, datagen, args)}), flush=True)
        return
    log(f"[bench] generating {n} uniform u32 keys (seed {SEED:#x})")
    keys = datagen.keys_u32(n, SEED, "uniform")
    info = N.plan_query(N.HPCS_DTYPE_U32, n)
    log(f"[bench] plan: {info.passes} passes x {info.digit_bits} bits, tile {info.tile_keys}, "
        f"{info.tiles} tiles, ideal {info.ideal_bytes / 1e9:.3f} GB")
    t_step, stats, t_events, ok = measure(torch, N, keys, args.steps, args.warmup, profile=True)
    if not ok:
        print("bench: sorted output differs from the reference sort", file=sys.stderr)
        sys.exit(3)
    gkeys = n / t_step / 1e9
    log(f"[bench] uniform: {t_step * 1e3:.3f} ms/sort, {gkeys:.2f} Gkeys/s; kernels {stats}")

    # dominant kernel: the one with the most time per sort among those that
    # move every key once (read + write: the pass scatter, onesweep, and in
    # the MSD-first form its bucket sort and its byte-2 scatter); its
    # algorithmic bytes per launch = one read and one write of every key
    movers = {"scatter": "k_onesweep as the pass scatter", "onesweep": "k_onesweep",
              "hyb_local": "k_hyb_local (the LDS bucket sort of the MSD-first form)",
              "seg_scatter": "k_seg_scatter (the byte-2 split of the MSD-first form)"}
    msd_first = "hyb_local" in stats
    dom = max((k for k in movers if k in stats), key=lambda k: stats[k][1])
    d_launch, d_ms = stats.get(dom, (0, 0.0))
    # launches that move keys: in the MSD-first form one of the LSD loop's
    # four scatter launches (three find nothing to do) and one byte-2 scatter
    real = {"scatter": args.steps, "seg_scatter": args.steps} if msd_first else {}
    d_avg_ms = d_ms / max(real.get(dom, d_launch), 1)
    d_bytes = n * 4 * 2
    d_gbs = d_bytes / (d_avg_ms * 1e-3) / 1e9 if d_avg_ms > 0 else 0.0
    sort_gbs = info.ideal_bytes / t_step / 1e9

    # HBM bytes per launch of the dominant kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes of this command (calibrated against a
    # 1 GiB copy); null when no profile of this workload is present.
    traffic = None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "sort_u32_2p28_uniform_traffic.json")
    if n == N_DEFAULT and os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        if tj.get("timed_name") == dom:  # a profile of this kernel
            traffic = int(tj["traffic_bytes_per_launch"])

    extra = {"sort_ms": round(t_step * 1e3, 4),
             "sort_ms_with_kernel_events": round(t_events * 1e3, 4),
             "sort_ideal_bytes": int(info.ideal_bytes),
             "sort_roofline_frac": round(sort_gbs / HBM_PEAK_GBS, 4),
             "kernel_ms_per_sort": {k: round(v[1] / args.steps, 4) for k, v in stats.items()},
             "kernel_launches_per_sort": {k: v[0] // args.steps for k, v in stats.items()},
             "sort_form": "MSD-first: top-byte and byte-2 splits, then 65536 LDS buckets (pass algorithm 6)"
                          if msd_first else "LSD passes"}
    if not args.no_dists:
        dists = {}
        for dist in ("low8", "sorted", "constant"):
            k2 = datagen.keys_u32(n, SEED + 1, dist)
            t2, _, _, ok2 = measure(torch, N, k2, max(2, args.steps // 2), 1, profile=False)
            if not ok2:
                print(f"bench: {dist} output differs from the reference sort", file=sys.stderr)
                sys.exit(3)
            dists[dist] = round(n / t2 / 1e9, 3)
            log(f"[bench] {dist}: {t2 * 1e3:.3f} ms/sort, {dists[dist]} Gkeys/s")
        extra["gkeys_by_distribution"] = dists

    if not args.no_batched:
        log("[bench] batched vs plain: 4096 segments")
        extra["batched_vs_plain"] = batched_vs_plain(torch, N, datagen, 4096, 0x5EED0008, 6)
        log(f"[bench] batched: {extra['batched_vs_plain']}")
        if not extra["batched_vs_plain"]["identical_outputs"]:
            print("bench: batched and plain outputs differ", file=sys.stderr)
            sys.exit(3)

    if not args.no_hist:
        extra["histogram"] = histogram_figures(torch, N, datagen, keys, 10)
        log(f"[bench] histograms and the 2^20 workload: {extra['histogram']}")
    if not args.no_host:
        extra["host_array"] = host_array_figures(N, keys, 2)
        log(f"[bench] host-array call: {extra['host_array']}")
    if not args.no_topk:
        extra["topk"] = topk_figures(torch, N, datagen, 5)
        log(f"[bench] top-k: {extra['topk']}")
    if not args.no_hist:
        extra["msd_split"] = msd_split_figures(torch, N, 10)
        log(f"[bench] MSD split (opt-in): {extra['msd_split']}")
    if not args.no_cfg3:
        extra["cfg3"] = cfg3_figures(torch, N, datagen, 3)
        if not args.no_cpu:
            extra["cfg3"]["cpu_baseline"] = cfg3_cpu_baseline(datagen)
        log(f"[bench] floating/record keys: {extra['cfg3']}")
    if not args.no_serving:
        log(f"[bench] serving: {args.serving_requests} requests")
        extra["serving"] = serving_figures(args.serving_requests, 4)
        log(f"[bench] serving: {extra['serving']}")

    cpu = None
    if not args.no_cpu:
        log(f"[bench] CPU baseline: SortRadixPar restatement on {n} keys")
        v, threads, times = cpu_baseline(keys, args.cpu_reps)
        cpu = {"value": round(v, 4), "unit": "Gkeys/s", "cores": threads, "kind": "port",
               "sample": f"full 2^{n.bit_length() - 1} uniform u32 array, 1 warm-up + median of {args.cpu_reps} "
                         f"(s: {', '.join(f'{t:.2f}' for t in times)}) on {cpu_model()}"}
        log(f"[bench] CPU: {v:.3f} Gkeys/s on {threads} threads")

    line = {
        "metric": "u32 LSD radix sort Gkeys/s @2^28 + %HBM roofline; batched vs plain",
        "value": round(gkeys, 3),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (splitmix64 uniform 32-bit keys, device-resident)",
        "config": {"workload": f"plain single-array stable LSD radix sort, uint32[2^{n.bit_length() - 1}] uniform",
                   "n": n, "digit_bits": info.digit_bits, "passes": info.passes, "tile_keys": info.tile_keys},
        "roofline": {"bound": "hbm", "achieved": round(d_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(d_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": d_bytes,
                     "kernel": f"{movers[dom]} (read + write 4 B per key); launch durations from HIP events "
                               f"on the launch stream over a second timed region of {args.steps} sorts"},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
