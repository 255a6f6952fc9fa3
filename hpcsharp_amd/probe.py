"""Short, fixed workloads for profiling runs (rocprofv3) of the top-k,
batched-segment and serving paths.

  python -m hpcsharp_amd.probe topk [--lg 26] [--k 10000] [--reps 5]
      hpcs_dev_topk_u32 then hpcs_dev_topk_u64 on 2^lg uniform keys,
      3 warm-up and `reps` timed calls each; prints ms per call.
  python -m hpcsharp_amd.probe segments [--nseg 4096] [--reps 5]
      hpcs_dev_sort_segments_u32 on nseg uint32 segments with lengths
      log-uniform in [16, 2^20] (seed 0x5EED0008), 1 warm-up and `reps` timed
      calls; prints ms per call.
  python -m hpcsharp_amd.probe histogram [--lg 28] [--reps 10]
      hpcs_dev_histogram_byte_components_u32 then _u64 on 2^lg uniform keys
      (the u32 array also ascending and constant), 3 warm-up and `reps` timed
      calls each, after one 1 GiB device copy (counter calibration); prints
      us per call.
  python -m hpcsharp_amd.probe serving [--requests 20000] [--slots 4]
      the serving workload (hpcsharp_amd.serving.make_trace) through
      hpcs_serve_run; prints the summary as JSON.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys


def topk(lg: int, k: int, reps: int) -> None:
    import torch
    from hpcsharp_amd import _native as N
    ctx, lib = N.context(0), N.lib()
    n = 1 << lg
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    g = torch.Generator(device="cuda").manual_seed(lg)
    for fn, words, dt in (("hpcs_dev_topk_u32", 1, torch.int32), ("hpcs_dev_topk_u64", 2, torch.int64)):
        x = torch.randint(-2 ** 31, 2 ** 31 - 1, (n * words,), dtype=torch.int32, device="cuda", generator=g).view(dt)
        out = torch.empty(k, dtype=dt, device="cuda")
        f = getattr(lib, fn)
        call = lambda: N.check(f(ctx, ctypes.c_void_p(x.data_ptr()), n, k, ctypes.c_void_p(out.data_ptr()), sp))
        for _ in range(3):
            call()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(reps):
            call()
        b.record(st)
        b.synchronize()
        print(f"{fn} n=2^{lg} k={k}: {a.elapsed_time(b) / reps:.4f} ms per call", flush=True)


def histogram(lg: int, reps: int) -> None:
    import torch
    from hpcsharp_amd import _native as N
    ctx, lib = N.context(0), N.lib()
    n = 1 << lg
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    g = torch.Generator(device="cuda").manual_seed(lg)
    # a 1 GiB device copy first: the known bytes that calibrate HBM counters
    src = torch.empty(1 << 28, dtype=torch.int32, device="cuda")
    torch.empty_like(src).copy_(src)
    del src
    cases = (("u32 uniform", 4, lambda: torch.randint(-2 ** 31, 2 ** 31 - 1, (n,), dtype=torch.int32, device="cuda",
                                                       generator=g)),
             ("u32 ascending", 4, lambda: torch.arange(n, dtype=torch.int32, device="cuda")),
             ("u32 constant", 4, lambda: torch.full((n,), 12345, dtype=torch.int32, device="cuda")),
             ("u64 uniform", 8, lambda: torch.randint(-2 ** 31, 2 ** 31 - 1, (2 * n,), dtype=torch.int32,
                                                       device="cuda", generator=g)))
    for name, kb, make in cases:
        x = make()
        counts = torch.empty(kb * 256, dtype=torch.int32, device="cuda")
        f = lib.hpcs_dev_histogram_byte_components_u32 if kb == 4 else lib.hpcs_dev_histogram_byte_components_u64
        call = lambda: N.check(f(ctx, ctypes.c_void_p(x.data_ptr()), n, ctypes.c_void_p(counts.data_ptr()), sp))
        for _ in range(3):
            call()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        for _ in range(reps):
            call()
        b.record(st)
        b.synchronize()
        print(f"histogram {name} n=2^{lg}: {a.elapsed_time(b) / reps * 1e3:.1f} us per call", flush=True)
        del x, counts


def segments(nseg: int, reps: int) -> None:
    import numpy as np
    import torch
    from hpcsharp_amd import _native as N
    from hpcsharp_amd import datagen
    seed = 0x5EED0008
    lens = datagen.segment_lengths(nseg, seed, 16, 1 << 20)
    begin = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(begin[-1])
    keys = datagen.keys_u32(total, seed + 1, "uniform")
    ctx, lib = N.context(0), N.lib()
    src = torch.from_numpy(keys.view(np.int32)).cuda()
    d_begin = torch.from_numpy(begin).cuda()
    work = src.clone()
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    call = lambda: N.check(lib.hpcs_dev_sort_segments_u32(ctx, ctypes.c_void_p(work.data_ptr()), total,
                                                          ctypes.c_void_p(d_begin.data_ptr()), nseg, sp))
    call()
    torch.cuda.synchronize()
    ms = []
    for _ in range(reps):
        work.copy_(src)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(st)
        call()
        b.record(st)
        b.synchronize()
        ms.append(a.elapsed_time(b))
    print(f"hpcs_dev_sort_segments_u32 {nseg} segments, {total} keys: {sorted(ms)[len(ms) // 2]:.4f} ms per call "
          f"(median of {reps})", flush=True)


def serving(requests: int, slots: int) -> None:
    from hpcsharp_amd import serving as S
    t = S.make_trace(nreq=requests)
    res, st = S.run(t, slots=slots)
    print(json.dumps(S.summarize(t, res, st)), flush=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("topk")
    a.add_argument("--lg", type=int, default=26)
    a.add_argument("--k", type=int, default=10000)
    a.add_argument("--reps", type=int, default=5)
    b = sub.add_parser("serving")
    b.add_argument("--requests", type=int, default=20000)
    b.add_argument("--slots", type=int, default=4)
    h = sub.add_parser("histogram")
    h.add_argument("--lg", type=int, default=28)
    h.add_argument("--reps", type=int, default=10)
    c = sub.add_parser("segments")
    c.add_argument("--nseg", type=int, default=4096)
    c.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    if args.what == "topk":
        topk(args.lg, args.k, args.reps)
    elif args.what == "histogram":
        histogram(args.lg, args.reps)
    elif args.what == "segments":
        segments(args.nseg, args.reps)
    else:
        serving(args.requests, args.slots)
    return 0


if __name__ == "__main__":
    sys.exit(main())
