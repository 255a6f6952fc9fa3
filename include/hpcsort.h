/*
 * hpcsort.h -- C ABI of libhpcsort.so, the MI355X (gfx950) implementation of
 * HPCsharp's stable LSD radix sort and byte-component histogram.
 *
 * Plain C: pointers, sizes and status codes only.  No HIP, C++ or torch
 * types cross this boundary; a stream is passed as `void*` (a hipStream_t,
 * NULL = HIP's null stream).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the HPCsharp source tree).  A .NET caller binds these entry points with
 * [LibraryImport] over pinned spans.
 *
 * Conventions shared by every entry point:
 *  - Return HPCS_OK (0) or a negative HPCS_E_* status; hpcs_last_error()
 *    returns a thread-local message for the last failure on this thread.
 *  - "Host-array" calls are synchronous and sort the caller's array in place
 *    (the reference's in-place interface: the result is in the caller's
 *    array, RadixSortLsd.cs:196-201, RadixSortLsdParallel.cs:223-231).
 *  - "hpcs_dev_*" calls take device pointers, are ordered on `stream` and
 *    return without synchronising.  Scratch memory belongs to the context;
 *    calls that share a context are serialised on the context's lock, and two
 *    streams must not use one context's scratch concurrently (use one context
 *    per stream for concurrent work).
 *  - Sorting is stable: equal keys keep their input order, so key-value and
 *    record sorts carry the exact stable permutation of the reference.
 *  - Floating-point keys use the reference's radix key (non-negative values
 *    flip the sign bit, negative values flip every bit, RadixSortMsd.cs:
 *    1379-1383, 1484-1488): ascending order with -0.0 before +0.0.
 *  - Element counts are int64_t; n must be below 2^31 elements per array, the
 *    reference's largest array (HPCS_E_UNSUPPORTED otherwise).
 */
#ifndef HPCSORT_H
#define HPCSORT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPCS_OK 0
#define HPCS_E_INVALID_ARG (-1) /* ArgumentNullException / ArgumentOutOfRangeException */
#define HPCS_E_OOM (-2)         /* OutOfMemoryException */
#define HPCS_E_HIP (-3)         /* a HIP runtime error (InvalidOperationException) */
#define HPCS_E_BUDGET (-4)      /* request exceeds the context's memory budget */
#define HPCS_E_UNSUPPORTED (-5) /* size or option outside this version */
#define HPCS_E_BUSY (-6)        /* resource in use */

#define HPCS_KEY_UINT 0  /* record key is an unsigned integer */
#define HPCS_KEY_FLOAT 1 /* record key is an IEEE float/double (radix key transform) */
#define HPCS_KEY_INT 2   /* two's-complement signed key (int / long) */

/* Tuning options for hpcs_ctx_set_option. */
#define HPCS_OPT_RANK 1       /* how a wave finds the keys that share its digit: 0: ballots, 1: LDS-OR match
                                 masks (default), 2 / 3: LDS-OR masks with every second / fourth item matched
                                 by ballots (the two methods load different pipes) */
#define HPCS_OPT_TILE_SHAPE 2 /* threads x keys per thread of a pass's tile: 0: 256x16 (default) and
                                 4: 256x8, any sort; 1: 512x16, 2: 256x32, 3: 512x32 and 5: 1024x16,
                                 4-byte keys without values; 6: 1024x8, 8-byte keys without values or
                                 4-byte keys with values.  Shapes 5 and 6 rank by ballot (HPCS_OPT_RANK
                                 is ignored for them).  A shape that does not serve a sort is ignored
                                 for it (the table's or the default shape is used) */
#define HPCS_OPT_FALLBACK_ROUNDS 3 /* look-back rounds on an unpublished tile before counting it locally
                                      (default 24; 0 makes every wait use the local count, for tests) */
#define HPCS_OPT_LOOKBACK_WINDOW 4 /* 4-byte keys without values, 256x16 tiles: earlier tiles read per look-back
                                      round: 16, 32 (default), 64; other sorts read 32 */
#define HPCS_OPT_PASS_ALGO 5       /* 0: single-pass onesweep with look-back, 1: reduce-then-scan,
                                      2: by size (default: one workgroup up to 16384 4-byte / 8192 8-byte
                                      keys, onesweep for small arrays, reduce-then-scan beyond),
                                      3: one workgroup sorts the array in LDS (larger arrays: by size),
                                      4: one MSD split by the top eight varying bits, then one workgroup per
                                      bucket (keys without values, up to 2^22 4-byte / 2^21 8-byte keys;
                                      other sorts: by size).  Opt-in only: for keys spread over their range
                                      it takes about half the time of 2: for 8-byte keys from 2^14 to 2^20,
                                      but keys clustered in one bucket (a bucket larger than a workgroup's
                                      tile is sorted by that one workgroup) cost 4-60x;
                                      5: runs of 4096 4-byte / 2048 8-byte keys sorted in LDS by one
                                      workgroup each, then stable merge passes (up to 2^20 keys; larger
                                      arrays: by size),
                                      6: 4-byte unsigned keys without values, 16385 to 2^28 + 2^25: MSD-first by
                                      the top byte, then byte 2, then each of the (up to 65536) buckets in
                                      LDS; keys whose top two bytes do not both vary take the passes of 1
                                      (float and signed keys always do; other sorts: by size) */
#define HPCS_OPT_DIGIT_BITS 8      /* 8 (default) or 11: 11-bit digits (3 passes for 4-byte keys, 6 for 8-byte
                                      keys) in the onesweep form, for sorts of keys without values below 2^30
                                      keys; other sorts keep 8-bit digits */
#define HPCS_OPT_DIGIT_BYTES 7     /* reduce-then-scan: 1 the scatter also writes each key's next digit as a
                                      byte and the next pass counts its tiles from those n bytes; 0 the next
                                      pass re-reads the keys; 2 (default) 1 for 8-byte keys with values, else 0
                                      (the n bytes are scratch only where they are written) */
#define HPCS_OPT_SCRATCH_LIMIT 12   /* bytes of device scratch one sort may use (0, the default: no limit).  A sort
                                      of keys without values whose scratch would exceed it, or whose scratch
                                      allocation fails (HPCS_E_OOM), sorts each half in place and merges the
                                      halves with half the keys of scratch (the GPU form of the reference's
                                      fallback to an in-place sort on OutOfMemoryException, RadixSortMsd.cs:
                                      725-770); beyond even that, and for sorts with values beyond the limit:
                                      HPCS_E_BUDGET, before any work.  Same output.  Applies to
                                      hpcs_sort_radix_* and hpcs_dev_sort_radix_* */
#define HPCS_OPT_GRAPHS 9          /* 1: graph cache for the device-resident sorts (hpcs_dev_sort_radix_*,
                                      hpcs_dev_sort_segments_u32) on a non-null stream: the first call with
                                      given pointers, size and stream runs directly, the second captures its
                                      launch sequence in a HIP graph, later ones launch that graph (one
                                      launch instead of up to ~30); 0 (default) plain launches.  Setting any
                                      option clears the cache; a call on a stream the caller is capturing
                                      launches plainly */
#define HPCS_OPT_SEG_RANK 13       /* hpcs_[dev_]sort_segments_u32: ranking of the large segments' scatter
                                      (as HPCS_OPT_RANK: 0 ballots, 1 LDS-OR, 2 / 3 LDS-OR with every second /
                                      fourth item by ballots); default from the tuning table, else 2 */
#define HPCS_OPT_SEG_SHAPE 15      /* hpcs_[dev_]sort_segments_u32: tile shape of the large segments' passes, a
                                      row of the shape table with 256 or 512 threads that serves 4-byte keys
                                      without values (0-4, 7-9; see HPCS_OPT_TILE_SHAPE); default from the
                                      tuning table, else 0 (256x16) */
#define HPCS_OPT_SEG_SMALL 14      /* hpcs_[dev_]sort_segments_u32: segments of at most this many keys are
                                      sorted by one workgroup in LDS (one read and one write per key): 4096 or
                                      16384; default from the tuning table, else 4096 */
#define HPCS_OPT_SEG_MSD 16        /* hpcs_[dev_]sort_segments_u32: segments of more than this many keys take
                                      the MSD-first form (one segmented pass by the top byte, then each of
                                      the 256 buckets sorted in LDS: 20 bytes of HBM traffic per key against
                                      48 for the four passes); same output.  0 off; otherwise at least
                                      16385; default 65536 */

#define HPCS_DTYPE_U32 0
#define HPCS_DTYPE_U64 1
#define HPCS_DTYPE_F32 2
#define HPCS_DTYPE_F64 3

typedef struct hpcs_ctx hpcs_ctx;

/* Sort plan for a dtype and size: what the sort moves, for roofline accounting. */
typedef struct hpcs_plan_info {
    int32_t digit_bits;      /* bits per radix digit */
    int32_t passes;          /* scatter passes in the full plan */
    int32_t tile_keys;       /* keys per onesweep tile (one workgroup) */
    int32_t block_threads;   /* threads per workgroup */
    int64_t tiles;           /* tiles per pass */
    int64_t ideal_bytes;     /* histogram read + (read + write) per pass, keys (+ values) */
    int64_t scratch_bytes;   /* device scratch the plan needs */
} hpcs_plan_info;

/* Per-kernel timing collected with HIP events recorded on the launch stream. */
typedef struct hpcs_kernel_stat {
    char name[32];     /* kernel name, e.g. "onesweep" */
    int64_t launches;  /* timed launches since profiling was enabled */
    double total_ms;   /* sum of event-measured durations */
} hpcs_kernel_stat;

/* ---- library and context ------------------------------------------------ */

int32_t hpcs_version(void);
const char *hpcs_last_error(void);
/* Create a context on HIP device `device`; owns scratch memory and a stream. */
int32_t hpcs_ctx_create(int32_t device, hpcs_ctx **out);
int32_t hpcs_ctx_destroy(hpcs_ctx *ctx);
/* Release the context's cached scratch memory. */
int32_t hpcs_ctx_trim(hpcs_ctx *ctx);
int32_t hpcs_plan_query(hpcs_ctx *ctx, int32_t dtype, int64_t n, int32_t with_values, hpcs_plan_info *info);
/* The tile (threads x keys per thread) that every pass kernel of a sort of
 * key_bytes (4 or 8) keys, with or without u32 values, runs with tile shape
 * `shape` (HPCS_OPT_TILE_SHAPE); HPCS_E_UNSUPPORTED for a shape that does not
 * serve that sort.  No device is needed.  No counterpart in the reference
 * (a diagnostic of this library's kernel dispatch). */
int32_t hpcs_tile_shape_query(int32_t shape, int32_t key_bytes, int32_t with_values, int32_t *block, int32_t *items);
/* Set a kernel tuning option (HPCS_OPT_*) for later calls on this context;
 * no counterpart in the reference (whose knobs are static properties such as
 * SortRadixParallelWorkQuanta, RadixSortLsdParallel.cs:763). */
int32_t hpcs_ctx_set_option(hpcs_ctx *ctx, int32_t option, int64_t value);
/* Grow the context's device scratch now, so that later sorts and top-k calls
 * of up to n keys of `dtype` allocate nothing (no counterpart in the
 * reference; for latency-bound callers such as the serving loop). */
int32_t hpcs_ctx_reserve(hpcs_ctx *ctx, int32_t dtype, int64_t n);
/* Size the context's scratch for hpcs_dev_sort_segments_u32 of up to n keys
 * in up to nseg segments, so those calls allocate nothing (no counterpart in
 * the reference: for serving loops that reserve memory up front). */
int32_t hpcs_ctx_reserve_segments(hpcs_ctx *ctx, int64_t n, int32_t nseg);
/* Tuning tables (no counterpart in the reference, whose tuning knobs are
 * static properties, RadixSortLsdParallel.cs:763, SystemAttributes.cs:9).
 * A table is text, one row per line (`#` comments):
 *   <dtype> <max_n> <algo> <shape> <rank> [<digit_bits>]
 * dtype u32|f32|i32|u64|f64|i64|pairs32|pairs64; a row applies to sorts of at
 * most max_n keys (the row with the smallest such max_n; beyond every row,
 * the largest); algo 0 onesweep, 1 reduce-then-scan, 3 one workgroup, 4 MSD
 * split + one workgroup per bucket (not for pairs rows; see
 * HPCS_OPT_PASS_ALGO); shape 0-4
 * (HPCS_OPT_TILE_SHAPE, a shape that serves the dtype; see
 * hpcs_tile_shape_query); rank 0-3 (HPCS_OPT_RANK; 0 for shapes 5 and 6);
 * digit_bits 8 (default) or 11 (HPCS_OPT_DIGIT_BITS; an
 * 11-bit row reads `<dtype> <max_n> 0 0 0 11`, keys without values only).
 * Options set with
 * hpcs_ctx_set_option (or HPCS_RANK / HPCS_TILE_SHAPE) take precedence.
 * hpcs_ctx_load_tuning replaces the context's table (NULL or "" clears it);
 * a context also loads $HPCS_TUNE_FILE when created.  hpcs_tuning_check
 * parses a table without a device and reports its row count. */
int32_t hpcs_ctx_load_tuning(hpcs_ctx *ctx, const char *path);
int32_t hpcs_tuning_check(const char *path, int32_t *rows);
/* Kernel timing (no counterpart in the reference, whose Stopwatch timing is
 * commented out, RadixSortLsd.cs:146-153): on != 0 clears the statistics and
 * brackets every later launch of this context with HIP events on its stream;
 * hpcs_profile_read waits for those events and returns up to max entries. */
int32_t hpcs_profile_enable(hpcs_ctx *ctx, int32_t on);
int32_t hpcs_profile_read(hpcs_ctx *ctx, hpcs_kernel_stat *out, int32_t max, int32_t *count);
/* Graph cache of HPCS_OPT_GRAPHS (no counterpart in the reference, whose
 * calls are synchronous CPU code; for callers that repeat the same
 * device-resident sort, such as a fixed-shape pipeline): graphs held now,
 * captures and graph launches since the context was created. */
int32_t hpcs_graph_cache_stats(hpcs_ctx *ctx, int64_t *entries, int64_t *captures, int64_t *replays);

/* ---- host-array drop-ins (synchronous, in place) ------------------------ */

/* Algorithm.SortRadixLsd(this uint[]) RadixSortLsd.cs:255-260 (and its
 * SortRadix compatibility name, HPCsharpExamples Program.cs:26);
 * ParallelAlgorithm.SortRadixPar(this uint[], int) RadixSortLsdParallel.cs:145-232. */
int32_t hpcs_sort_radix_u32(hpcs_ctx *ctx, uint32_t *keys, int64_t n);
/* Algorithm.SortRadixLsd(this uint[], int startIndex, int length) with the
 * Experimental overload's correct sub-range semantics, RadixSortLsd.cs:1409-1457. */
int32_t hpcs_sort_radix_range_u32(hpcs_ctx *ctx, uint32_t *keys, int64_t start, int64_t length);
/* Algorithm.SortRadixLsd(this ulong[]) RadixSortLsd.cs:359-442;
 * ParallelAlgorithm.SortRadixPar(this ulong[]) RadixSortLsdParallel.cs:690-734. */
int32_t hpcs_sort_radix_u64(hpcs_ctx *ctx, uint64_t *keys, int64_t n);
/* Algorithm.SortRadixMsd(this float[]) RadixSortMsd.cs:1563-1571 (float order). */
int32_t hpcs_sort_radix_f32(hpcs_ctx *ctx, float *keys, int64_t n);
/* Algorithm.SortRadixMsd(this double[]) RadixSortMsd.cs:1587-1595 (double order). */
int32_t hpcs_sort_radix_f64(hpcs_ctx *ctx, double *keys, int64_t n);
/* Algorithm.SortRadix<T>(T[] items, UInt64[] keys) RadixSortLsd.cs:1245-1299,
 * full 64-bit key, uint32 items; both arrays are permuted in place. */
int32_t hpcs_sort_radix_pairs_u64_u32(hpcs_ctx *ctx, uint64_t *keys, uint32_t *vals, int64_t n);
/* Algorithm.SortRadix<T>(T[] items, UInt32[] keys) RadixSortLsd.cs:1183-1237. */
int32_t hpcs_sort_radix_pairs_u32_u32(hpcs_ctx *ctx, uint32_t *keys, uint32_t *vals, int64_t n);
/* Key-extractor record sorts: SortRadixLsd2<T>(Func<T,UInt64>) RadixSortLsd.cs:
 * 821-846 (key_bits 64), SortRadixLsd2<T>(Func<T,UInt32>) :751-775 (key_bits 32),
 * for unmanaged records of rec_size bytes whose key sits at key_offset. */
int32_t hpcs_sort_radix_records(hpcs_ctx *ctx, void *recs, int64_t n, int32_t rec_size, int32_t key_offset,
                                int32_t key_bits, int32_t key_kind);
/* Algorithm.HistogramByteComponents(uint[], int l, int r) Histogram.cs:261-296;
 * ParallelAlgorithm.HistogramByteComponentsPar HistogramParallel.cs:214-222.
 * l..r inclusive; counts is int32[4][256], digit-major. */
int32_t hpcs_histogram_byte_components_u32(hpcs_ctx *ctx, const uint32_t *a, int64_t l, int64_t r, int32_t *counts);
/* Algorithm.HistogramByteComponents(ulong[], int l, int r) Histogram.cs:957-990;
 * HistogramByteComponentsPar(ulong[]) HistogramParallel.cs:545-555.  counts is int32[8][256]. */
int32_t hpcs_histogram_byte_components_u64(hpcs_ctx *ctx, const uint64_t *a, int64_t l, int64_t r, int32_t *counts);

/* A batch of uint32 segments, each sorted in place: segment s is
 * keys[seg_begin[s], seg_begin[s+1]) (nseg + 1 non-decreasing offsets);
 * keys outside every segment are untouched.  Replaces one call per segment of
 * HPCsharpExperimental.Algorithm.SortRadixLsd(uint[], startIndex, length),
 * RadixSortLsd.cs:1409-1457; segments must lie below 2^32 keys. */
int32_t hpcs_sort_segments_u32(hpcs_ctx *ctx, uint32_t *keys, const int64_t *seg_begin, int32_t nseg);

/* ---- device-resident, stream-ordered ------------------------------------ */

/* The same sorts on device buffers (d_* are device pointers). */
int32_t hpcs_dev_sort_radix_u32(hpcs_ctx *ctx, uint32_t *d_keys, int64_t n, void *stream);
/* hpcs_dev_sort_radix_u32 that allocates nothing: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch is a static device array of 26 MiB in the library (one per GPU,
 * held until the process ends), so n is at most 2^22 keys
 * (HPCS_E_UNSUPPORTED beyond).  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other contexts or streams) race, so the
 * caller orders them.  Options and the tuning table apply as for
 * hpcs_dev_sort_radix_u32, with 8-bit digits and no MSD split; profiling and
 * the graph cache do not record this call.  Same output, bit for bit. */
int32_t hpcs_dev_sort_radix_u32_noalloc(hpcs_ctx *ctx, uint32_t *d_keys, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_u64(hpcs_ctx *ctx, uint64_t *d_keys, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_f32(hpcs_ctx *ctx, float *d_keys, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_f64(hpcs_ctx *ctx, double *d_keys, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_pairs_u64_u32(hpcs_ctx *ctx, uint64_t *d_keys, uint32_t *d_vals, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_pairs_u32_u32(hpcs_ctx *ctx, uint32_t *d_keys, uint32_t *d_vals, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_records(hpcs_ctx *ctx, void *d_recs, int64_t n, int32_t rec_size, int32_t key_offset,
                                    int32_t key_bits, int32_t key_kind, void *stream);
/* Batched segments on the device: d_seg_begin is a device array of nseg + 1
 * non-decreasing offsets into d_keys[0, n).  The offsets are checked on the
 * device: a segment outside [0, n), with end < begin, or starting before the
 * end of an earlier segment is skipped (its keys untouched).  The launch
 * sequence does not depend on the segment lengths (planning runs on the
 * device), so after one warm-up call (which sizes the scratch) the call can
 * be captured in a HIP graph. */
int32_t hpcs_dev_sort_segments_u32(hpcs_ctx *ctx, uint32_t *d_keys, int64_t n, const int64_t *d_seg_begin,
                                   int32_t nseg, void *stream);
/* Byte-component histogram of d_keys[0, n): d_counts is uint32[4][256]
 * (u32) or uint32[8][256] (u64) on the device, overwritten. */
int32_t hpcs_dev_histogram_byte_components_u32(hpcs_ctx *ctx, const uint32_t *d_keys, int64_t n, uint32_t *d_counts,
                                               void *stream);
int32_t hpcs_dev_histogram_byte_components_u64(hpcs_ctx *ctx, const uint64_t *d_keys, int64_t n, uint32_t *d_counts,
                                               void *stream);

/* Replace SortRadix<T>(this T[] inputArray, UInt32[] / UInt64[] inKeys)
 * (RadixSortLsd.cs:1183-1299) for items of any fixed size: keys (key_bits 32
 * or 64, unsigned) are sorted stably in place and items[i] (item_size bytes
 * each) moves with keys[i].  The full key is sorted (the reference's UInt64
 * overload stops after four passes, :1258, :1284). */
int32_t hpcs_sort_radix_keys_items(hpcs_ctx *ctx, void *keys, int32_t key_bits, void *items, int32_t item_size,
                                   int64_t n);
int32_t hpcs_dev_sort_radix_keys_items(hpcs_ctx *ctx, void *d_keys, int32_t key_bits, void *d_items,
                                       int32_t item_size, int64_t n, void *stream);

/* Replace Histogram11bitComponents(uint[], l, r) (Histogram.cs:298-331; bits
 * = 11) and HistogramNbitComponents(uint[] / ulong[], l, r, bits)
 * (Histogram.cs:381-427): digit d is (key >> (d * bits)) & (2^bits - 1), d <
 * ceil(key width / bits); counts is int32[digits][2^bits] row-major; 1 <=
 * bits <= 16. */
int32_t hpcs_histogram_nbit_components_u32(hpcs_ctx *ctx, const uint32_t *a, int64_t l, int64_t r, int32_t bits,
                                           int32_t *counts);
int32_t hpcs_histogram_nbit_components_u64(hpcs_ctx *ctx, const uint64_t *a, int64_t l, int64_t r, int32_t bits,
                                           int32_t *counts);
int32_t hpcs_dev_histogram_nbit_components_u32(hpcs_ctx *ctx, const uint32_t *d_keys, int64_t n, int32_t bits,
                                               uint32_t *d_counts, void *stream);
int32_t hpcs_dev_histogram_nbit_components_u64(hpcs_ctx *ctx, const uint64_t *d_keys, int64_t n, int32_t bits,
                                               uint32_t *d_counts, void *stream);
/* Replace HistogramByteComponents<T>(T[] inArray, int l, int r,
 * Func<T, UInt32> / Func<T, UInt64> getKey) (Histogram.cs:760-844) for a
 * key stored at key_offset inside records of rec_size bytes: counts is
 * int32[key_bits / 8][256] of the key bytes of records l..r (inclusive). */
int32_t hpcs_histogram_byte_components_records(hpcs_ctx *ctx, const void *recs, int64_t l, int64_t r, int32_t rec_size,
                                               int32_t key_offset, int32_t key_bits, int32_t *counts);
int32_t hpcs_dev_histogram_byte_components_records(hpcs_ctx *ctx, const void *d_recs, int64_t n, int32_t rec_size,
                                                   int32_t key_offset, int32_t key_bits, uint32_t *d_counts,
                                                   void *stream);

/* ------------------------------------------------------------------------
 * Signed keys: int[] and long[]
 * ------------------------------------------------------------------------ */

/* Replace Algorithm.SortRadixLsd(this int[]) / (this long[])
 * (RadixSortLsd.cs:267-357, 444-532) and ParallelAlgorithm.SortRadixPar(int[]
 * / long[]) (RadixSortLsdParallel.cs:446-668): ascending, in place. */
int32_t hpcs_sort_radix_i32(hpcs_ctx *ctx, int32_t *keys, int64_t n);
int32_t hpcs_sort_radix_i64(hpcs_ctx *ctx, int64_t *keys, int64_t n);
int32_t hpcs_dev_sort_radix_i32(hpcs_ctx *ctx, int32_t *d_keys, int64_t n, void *stream);
int32_t hpcs_dev_sort_radix_i64(hpcs_ctx *ctx, int64_t *d_keys, int64_t n, void *stream);
/* Replace Algorithm.HistogramByteComponents(int[] / long[], l, r)
 * (Histogram.cs:992-1052): as the unsigned histograms, except that the top
 * digit counts (byte ^ 128), the reference's order of signed keys. */
int32_t hpcs_histogram_byte_components_i32(hpcs_ctx *ctx, const int32_t *a, int64_t l, int64_t r, int32_t *counts);
int32_t hpcs_histogram_byte_components_i64(hpcs_ctx *ctx, const int64_t *a, int64_t l, int64_t r, int32_t *counts);
int32_t hpcs_dev_histogram_byte_components_i32(hpcs_ctx *ctx, const int32_t *d_keys, int64_t n, uint32_t *d_counts,
                                               void *stream);
int32_t hpcs_dev_histogram_byte_components_i64(hpcs_ctx *ctx, const int64_t *d_keys, int64_t n, uint32_t *d_counts,
                                               void *stream);

/* ------------------------------------------------------------------------
 * Radix top-k
 * ------------------------------------------------------------------------ */

/* Replaces Algorithm.SelectTopK(this uint[] a, int start, int length, int k)
 * (SelectionTopKRadix.cs:142-153; with start = 0 and length = a.Length the
 * overload SelectTopK(this uint[] a, int k), :160-171).  k is an index into
 * a, start <= k < start + length.  On return a[k, start + length) holds the
 * largest start + length - k keys of the range in ascending order (so a[k] is
 * the k-th smallest of the range counted from start), and a[start, k) the
 * other keys in unspecified order.  The reference guarantees the multiset of
 * a[k, start + length) only, and leaves a[start, k) holding moved-from
 * copies; this call keeps the range a permutation of its input. */
int32_t hpcs_select_topk_u32(hpcs_ctx *ctx, uint32_t *a, int64_t start, int64_t length, int64_t k);

/* The k largest keys of keys[0, n) in descending order into out[0, k),
 * 0 <= k <= n; keys is not modified.  Floating keys order as the radix sort
 * orders them (-0.0 below +0.0).  Host arrays, synchronous. */
int32_t hpcs_topk_u32(hpcs_ctx *ctx, const uint32_t *keys, int64_t n, int64_t k, uint32_t *out);
int32_t hpcs_topk_u64(hpcs_ctx *ctx, const uint64_t *keys, int64_t n, int64_t k, uint64_t *out);
int32_t hpcs_topk_f32(hpcs_ctx *ctx, const float *keys, int64_t n, int64_t k, float *out);
int32_t hpcs_topk_f64(hpcs_ctx *ctx, const double *keys, int64_t n, int64_t k, double *out);
int32_t hpcs_topk_i32(hpcs_ctx *ctx, const int32_t *keys, int64_t n, int64_t k, int32_t *out);
int32_t hpcs_topk_i64(hpcs_ctx *ctx, const int64_t *keys, int64_t n, int64_t k, int64_t *out);
/* The same on device buffers, stream-ordered; the launch sequence depends
 * only on (n, k), so the call can be captured in a HIP graph after one
 * warm-up call of the same size. */
int32_t hpcs_dev_topk_u32(hpcs_ctx *ctx, const uint32_t *d_keys, int64_t n, int64_t k, uint32_t *d_out,
                          void *stream);
int32_t hpcs_dev_topk_u64(hpcs_ctx *ctx, const uint64_t *d_keys, int64_t n, int64_t k, uint64_t *d_out,
                          void *stream);
int32_t hpcs_dev_topk_f32(hpcs_ctx *ctx, const float *d_keys, int64_t n, int64_t k, float *d_out, void *stream);
int32_t hpcs_dev_topk_f64(hpcs_ctx *ctx, const double *d_keys, int64_t n, int64_t k, double *d_out, void *stream);
int32_t hpcs_dev_topk_i32(hpcs_ctx *ctx, const int32_t *d_keys, int64_t n, int64_t k, int32_t *d_out,
                          void *stream);
int32_t hpcs_dev_topk_i64(hpcs_ctx *ctx, const int64_t *d_keys, int64_t n, int64_t k, int64_t *d_out,
                          void *stream);

/* ------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference: the serving pattern built
 * on the sort and top-k calls above)
 * ------------------------------------------------------------------------ */

#define HPCS_OP_SORT 0 /* ascending sort of the payload (radix order) */
#define HPCS_OP_TOPK 1 /* the k largest keys of the payload, descending */

#define HPCS_SERVE_REALTIME 0 /* replay the arrival times; slots run concurrently */
#define HPCS_SERVE_ALONE 1    /* one request at a time, in list order */

/* One request.  Its payload is n keys of `dtype`, key i being output i of
 * splitmix64(seed) (for floating types the same bits with an all-ones
 * exponent made finite); it is generated on the device at admission. */
typedef struct hpcs_request {
    int32_t op;       /* HPCS_OP_SORT or HPCS_OP_TOPK */
    int32_t dtype;    /* HPCS_DTYPE_* */
    int64_t n;        /* payload keys, 1 <= n < 2^31 */
    int64_t k;        /* top-k: 1 <= k <= n */
    uint64_t seed;    /* payload seed */
    double arrival_s; /* arrival, seconds after the start (non-decreasing) */
} hpcs_request;

typedef struct hpcs_request_result {
    double admit_s;    /* slot and memory granted */
    double enqueued_s; /* all of the request's work enqueued on its slot's stream */
    double done_s;     /* completion seen by the host */
    uint64_t checksum; /* sum over i of mix64(bits_i + (i + 1) * 0x9E3779B97F4A7C15)
                          of the result keys, mod 2^64 (mix64: the splitmix64
                          finaliser) */
    int32_t status;    /* HPCS_OK, or HPCS_E_BUDGET if the request alone exceeds the budget */
    int32_t slot;      /* worker slot that ran it */
} hpcs_request_result;

typedef struct hpcs_serve_stats {
    double wall_s;                  /* start to the last completion */
    int64_t keys;                   /* payload keys of the completed requests */
    int64_t peak_bytes;             /* most request bytes in flight at once */
    int64_t request_budget_bytes;   /* budget left for requests after the slots' scratch */
    int64_t scratch_bytes_per_slot; /* scratch reserved per slot */
    int32_t slots;
    int32_t pad;
    int64_t batches;                /* launch chains enqueued: single requests plus coalesced batches */
    int64_t coalesced_requests;     /* requests that ran in a coalesced batch of two or more */
} hpcs_serve_stats;

typedef struct hpcs_serve_config {
    int32_t slots;             /* worker slots, 1-64 (one stream and one context each; HPCS_SERVE_ALONE: 1) */
    int32_t mode;              /* HPCS_SERVE_REALTIME or HPCS_SERVE_ALONE */
    int64_t hbm_budget;        /* bytes of device memory: request buffers plus the slots' scratch */
    double request_timeout_s;  /* longest wait for a request's (or batch's) completion once its slot's
                                  earlier requests are done; past it the run fails with HPCS_E_HIP and a
                                  message naming the request.  0: 5 s + 20 ns per key */
    int64_t coalesce_max_keys; /* sorts of u32 / f32 payloads of at most this many keys that are queued
                                  together are admitted as one batch and run as one segmented sort
                                  (hpcs_dev_sort_segments_u32, floats in their radix order); same results.
                                  0: no coalescing */
    int32_t coalesce_max_requests; /* requests per batch, 2-1024 (0: 256) */
    int32_t pad;
    int64_t coalesce_max_batch_keys; /* keys per batch (0: 2^22) */
} hpcs_serve_config;

/* Run the requests on `device` under `cfg`.  Admission is first come, first
 * served; a request that cannot fit beside the slots' scratch on its own
 * gets status HPCS_E_BUDGET.  results[i] belongs to reqs[i]; on a failed
 * run (the return code, hpcs_serve_last_error naming the first error) the
 * results are not meaningful. */
int32_t hpcs_serve_run_config(int32_t device, const hpcs_serve_config *cfg, const hpcs_request *reqs, int64_t nreq,
                              hpcs_request_result *results, hpcs_serve_stats *stats);
/* hpcs_serve_run_config with `slots`, `mode` and hbm_budget, the derived
 * request timeout and no coalescing. */
int32_t hpcs_serve_run(int32_t device, int32_t slots, int64_t hbm_budget, const hpcs_request *reqs, int64_t nreq,
                       int32_t mode, hpcs_request_result *results, hpcs_serve_stats *stats);
const char *hpcs_serve_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HPCSORT_H */
