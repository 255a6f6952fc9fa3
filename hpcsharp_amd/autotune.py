"""Autotuner: measure every sort variant per key type and size class on this
GPU and write the tuning table libhpcsort.so reads (hpcs_ctx_load_tuning).

Variants: digit width (8 bits; 11 bits for keys without values), pass
algorithm (onesweep with look-back, reduce-then-scan, one workgroup in LDS
for small arrays, block-sort + merge up to 2^20 keys, the MSD-first form for
4-byte keys without values), tile shape (every row of the library's shape table that
serves the key type, hpcs_tile_shape_query: 256, 512 and 1024 threads, 8 to
32 keys per thread) and ranking (ballot, LDS-OR, LDS-OR with every second or
fourth item by ballot; ballot only on the 1024-thread shapes).
Size classes are powers of two; a class's row applies to every n up to it.
The batched segmented sort (kind "segs") is tuned per class of total keys:
the largest segment one workgroup sorts in LDS (4096 or 16384), the tile
shape of the large segments' passes (the 256- and 512-thread rows, 8 to 32
keys per thread) and the ranking of their scatter, on a batch of segments
with log-uniform lengths in [16, min(2^20, total / 16)].
Each measured cell is appended to a JSON log as it finishes, and a rerun
skips the cells the log already holds (the sweep resumes).

Run on the GPU:  python -m hpcsharp_amd.autotune [--classes 12,13,14,16,18,20,22,24,26,28]
"""
from __future__ import annotations

import argparse
import ctypes
import itertools
import json
import os
import sys
import time

from hpcsharp_amd import _native as N

HERE = os.path.dirname(os.path.abspath(__file__))
# (table dtype, key bytes, with u32 values, entry point)
KINDS = {
    "u32": (4, False, "hpcs_dev_sort_radix_u32"),
    "u64": (8, False, "hpcs_dev_sort_radix_u64"),
    "pairs32": (4, True, "hpcs_dev_sort_radix_pairs_u32_u32"),
    "pairs64": (8, True, "hpcs_dev_sort_radix_pairs_u64_u32"),
}
RANKS = (1, 2, 3, 0)  # LDS-OR, every 2nd / 4th item by ballot, ballot
SEG_SMALL = (4096, 16384)  # largest segment sorted in LDS (HPCS_OPT_SEG_SMALL)
SEG_SHAPES = (0, 1, 2, 3, 4, 7, 8, 9)  # tile shapes of the large segments (HPCS_OPT_SEG_SHAPE)


# largest arrays the one-workgroup sort (pass algorithm 3) and the
# block-sort + merge form (pass algorithm 5) take
BLOCK_SORT_MAX = {4: 16384, 8: 8192}
MERGE_SORT_MAX = 1 << 20


def candidates(key_bytes: int, vals: bool, n: int | None = None):
    """8-bit digits: every pass algorithm x tile shape x ranking, and the
    one-workgroup sort where n fits it, the block-sort + merge form from
    there up to 2^20 keys (not the MSD split, pass algorithm 4:
    its speed depends on how the keys spread, which uniform keys do not
    show); 11-bit digits (keys without values):
    their one form, onesweep passes with 256x16 tiles and ballot ranking."""
    out = [{"algo": a, "shape": s, "rank": r, "bits": 8}
           for a, (s, ballot_only), r in itertools.product((0, 1), shapes_for(key_bytes, vals), RANKS)
           if r == 0 or not ballot_only]
    if not vals:
        out.append({"algo": 0, "shape": 0, "rank": 0, "bits": 11})
    if n is not None and n <= BLOCK_SORT_MAX[key_bytes]:
        out.append({"algo": 3, "shape": 0, "rank": 1, "bits": 8})
    if n is not None and BLOCK_SORT_MAX[key_bytes] < n <= MERGE_SORT_MAX:
        out.append({"algo": 5, "shape": 0, "rank": 1, "bits": 8})
    if key_bytes == 4 and not vals and n is not None and n > BLOCK_SORT_MAX[4]:
        # the MSD-first form: its top-byte pass and its segmented byte-2 pass
        # take the shape (rows of at most 512 threads) and the ranking
        out.extend({"algo": 6, "shape": s, "rank": r, "bits": 8}
                   for (s, ballot_only), r in itertools.product(shapes_for(4, False), RANKS)
                   if not ballot_only)
    return out


def shapes_for(key_bytes: int, vals: bool):
    """(shape, ballot only) for every row of the library's tile-shape table
    that serves the key type (the kernel dispatch answers, no device)."""
    out = []
    for shape in range(64):
        b, i = ctypes.c_int32(0), ctypes.c_int32(0)
        rc = N.lib().hpcs_tile_shape_query(shape, key_bytes, int(vals), ctypes.byref(b), ctypes.byref(i))
        if rc == N.HPCS_OK:
            out.append((shape, b.value >= 1024))
        elif rc != N.HPCS_E_UNSUPPORTED:
            N.check(rc)
    return out


def cell_key(kind: str, lg: int, cand: dict) -> str:
    key = f"{kind}:{lg}:{cand['algo']}:{cand['shape']}:{cand['rank']}"
    return key + ":11" if cand.get("bits", 8) == 11 else key


def time_variant(torch, ctx, kind: str, lg: int, cand: dict, reps: int) -> float:
    """Median seconds per sort of 2^lg uniform keys (device-resident)."""
    key_bytes, vals, fn = KINDS[kind]
    n = 1 << lg
    lib = N.lib()
    N.set_option(N.HPCS_OPT_PASS_ALGO, cand["algo"], ctx)
    N.set_option(N.HPCS_OPT_TILE_SHAPE, cand["shape"], ctx)
    N.set_option(N.HPCS_OPT_RANK, cand["rank"], ctx)
    N.set_option(N.HPCS_OPT_DIGIT_BITS, cand.get("bits", 8), ctx)
    g = torch.Generator(device="cuda").manual_seed(lg * 1000 + key_bytes)
    src = torch.randint(-2 ** 31, 2 ** 31 - 1, (n * key_bytes // 4,), dtype=torch.int32, device="cuda", generator=g)
    keys = src.clone()
    v = torch.arange(n, dtype=torch.int32, device="cuda") if vals else None
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    f = getattr(lib, fn)

    def call():
        if vals:
            N.check(f(ctx, ctypes.c_void_p(keys.data_ptr()), ctypes.c_void_p(v.data_ptr()), n, sp))
        else:
            N.check(f(ctx, ctypes.c_void_p(keys.data_ptr()), n, sp))
    times = []
    for r in range(reps + 1):
        keys.copy_(src)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(st)
        call()
        end.record(st)
        end.synchronize()
        if r:
            times.append(start.elapsed_time(end) * 1e-3)
    times.sort()
    return times[len(times) // 2]


def seg_batch(lg: int):
    """Segment lengths (log-uniform in [16, min(2^20, 2^lg / 16)]) whose
    expected total is 2^lg keys."""
    import numpy as np
    from hpcsharp_amd import datagen
    lo, hi = 16, min(1 << 20, (1 << lg) // 16)
    mean = (hi - lo) / np.log(hi / lo)
    nseg = max(1, int(round((1 << lg) / mean)))
    return datagen.segment_lengths(nseg, 0x5E6 + lg, lo, hi)


def time_segments(torch, ctx, lg: int, small: int, rank: int, shape: int, reps: int) -> float:
    """Median seconds per batched call (device-resident keys)."""
    import numpy as np
    N.set_option(N.HPCS_OPT_SEG_SMALL, small, ctx)
    N.set_option(N.HPCS_OPT_SEG_RANK, rank, ctx)
    N.set_option(N.HPCS_OPT_SEG_SHAPE, shape, ctx)
    lens = seg_batch(lg)
    begin = torch.from_numpy(np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)).cuda()
    n = int(lens.sum())
    g = torch.Generator(device="cuda").manual_seed(lg)
    src = torch.randint(-2 ** 31, 2 ** 31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    keys = src.clone()
    st = torch.cuda.current_stream()
    sp = ctypes.c_void_p(st.cuda_stream)
    times = []
    for r in range(reps + 1):
        keys.copy_(src)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(st)
        N.check(N.lib().hpcs_dev_sort_segments_u32(ctx, ctypes.c_void_p(keys.data_ptr()), n,
                                                   ctypes.c_void_p(begin.data_ptr()), int(lens.size), sp))
        end.record(st)
        end.synchronize()
        if r:
            times.append(start.elapsed_time(end) * 1e-3)
    times.sort()
    return times[len(times) // 2]


def table_text(results: dict) -> str:
    lines = ["# libhpcsort tuning table (hpcs_ctx_load_tuning), measured on "
             + results.get("device", "this GPU") + " by hpcsharp_amd.autotune",
             "# dtype max_n algo shape rank digit_bits   # median ms per sort of the best variant"]
    alias = {"u32": ("u32", "f32", "i32"), "u64": ("u64", "f64", "i64"), "pairs32": ("pairs32",),
             "pairs64": ("pairs64",)}
    for kind in KINDS:
        for lg in sorted({int(k.split(":")[1]) for k in results["cells"] if k.startswith(kind + ":")}):
            cells = [c for k, c in results["cells"].items() if k.startswith(f"{kind}:{lg}:")]
            best = min(cells, key=lambda c: c["s"])
            for name in alias[kind]:
                lines.append(f"{name} {1 << lg} {best['algo']} {best['shape']} {best['rank']} {best.get('bits', 8)}"
                             f"   # {best['s'] * 1e3:.4f} ms")
    lines.append("# segs max_n small_max rank shape   # median ms per batched call of the best variant")
    for lg in sorted({int(k.split(":")[1]) for k in results["cells"] if k.startswith("segs:")}):
        cells = [c for k, c in results["cells"].items() if k.startswith(f"segs:{lg}:")]
        best = min(cells, key=lambda c: c["s"])
        lines.append(f"segs {1 << lg} {best['small']} {best['rank']} {best.get('shape', 0)}"
                     f"   # {best['s'] * 1e3:.4f} ms")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--classes", default="12,13,14,16,18,20,22,24,26,28")
    ap.add_argument("--kinds", default="u32,u64,pairs32,pairs64,segs")
    ap.add_argument("--seg-classes", default="16,20,24,28")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--log", default=os.path.join(HERE, "tune_gfx950.json"))
    ap.add_argument("--out", default=N.TUNE_PATH)
    args = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        print("autotune needs a HIP device", file=sys.stderr)
        return 1
    prop = torch.cuda.get_device_properties(0)
    results = {"device": f"{prop.name} ({getattr(prop, 'gcnArchName', 'unknown arch')})", "cells": {}}
    if os.path.exists(args.log):
        with open(args.log) as f:
            results = json.load(f)
    ctx = N.new_context(0)
    try:
        for kind in args.kinds.split(","):
            if kind == "segs":
                for lg in (int(x) for x in args.seg_classes.split(",")):
                    for small, rank, shape in itertools.product(SEG_SMALL, RANKS, SEG_SHAPES):
                        key = f"segs:{lg}:{small}:{rank}:{shape}"
                        if key in results["cells"]:
                            continue
                        s = time_segments(torch, ctx, lg, small, rank, shape, args.reps)
                        results["cells"][key] = {"small": small, "rank": rank, "shape": shape, "s": s}
                        print(f"[autotune] {key}: {s * 1e3:.4f} ms", flush=True)
                        with open(args.log, "w") as f:
                            json.dump(results, f, indent=1, sort_keys=True)
                continue
            key_bytes, vals, _ = KINDS[kind]
            for lg in (int(x) for x in args.classes.split(",")):
                for cand in candidates(key_bytes, vals, 1 << lg):
                    key = cell_key(kind, lg, cand)
                    if key in results["cells"]:
                        continue                     # measured in an earlier run
                    s = time_variant(torch, ctx, kind, lg, cand, args.reps)
                    results["cells"][key] = dict(cand, s=s)
                    print(f"[autotune] {key}: {s * 1e3:.4f} ms", flush=True)
                    with open(args.log, "w") as f:
                        json.dump(results, f, indent=1, sort_keys=True)
    finally:
        N.destroy_context(ctx)
    with open(args.out, "w") as f:
        f.write(table_text(results))
    print(f"[autotune] wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
