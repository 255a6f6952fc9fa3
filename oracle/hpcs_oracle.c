/*
 * hpcs_oracle.c -- CPU restatement of HPCsharp's LSD radix sort path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests under
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py are the
 * only code that may load it.  The product (libhpcsort.so) never calls it and
 * never falls back to it.
 *
 * Every function restates one reference function; the reference file:line it
 * follows is cited above it.  Paths are relative to the HPCsharp source tree.
 * The reference is C# (.NET 7-9); no .NET toolchain exists here, so the
 * reference cannot be compiled or run (no oracle/_ref build).  Each function is
 * pinned in tests/test_oracle.py by an independent check (closed form, brute
 * force, the reference's own example vectors, invariants).
 *
 * Index convention: int64_t lengths (the reference uses int, N < 2^31).
 * Histogram ranges are inclusive l..r, as in the reference.
 *
 * Decisions on reference quirks (the contract is the correct behaviour):
 *  - SortRadixLsd(uint[], start, len) in RadixSortLsd.cs:127-202 bases its bin
 *    offsets at 0 (:120) while reading [start, end] (:165); that is wrong for
 *    start > 0.  The restatement bases the offsets at `start`, as the
 *    Experimental overload does (RadixSortLsd.cs:1436).
 *  - The 64-bit key-extractor overloads that loop on a 32-bit mask
 *    (RadixSortLsd.cs:795,814) sort only the low 32 bits.  Records here sort
 *    by the full key width, as SortRadixLsd2<T>(Func<T,UInt64>) does
 *    (RadixSortLsd.cs:821-846); key_bits = 32 selects a 32-bit key.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define HPCS_ORACLE_OK 0
#define HPCS_ORACLE_EINVAL (-1)
#define HPCS_ORACLE_ENOMEM (-2)

/* ------------------------------------------------------------------------ */
/* Byte-component histograms                                                 */
/* ------------------------------------------------------------------------ */

/* Algorithm.HistogramByteComponents(uint[] inArray, int l, int r)
 * Histogram.cs:261-296 -- one read of [l, r], four counters per key.
 * count is [4][256] row-major (digit-major). */
int hpcs_oracle_histogram_byte_components_u32(const uint32_t *a, int64_t l, int64_t r, int32_t *count)
{
    if (!a && r >= l) return HPCS_ORACLE_EINVAL;
    memset(count, 0, sizeof(int32_t) * 4 * 256);
    for (int64_t i = l; i <= r; i++) {
        uint32_t v = a[i];
        count[0 * 256 + (v & 0xff)]++;
        count[1 * 256 + ((v >> 8) & 0xff)]++;
        count[2 * 256 + ((v >> 16) & 0xff)]++;
        count[3 * 256 + ((v >> 24) & 0xff)]++;
    }
    return HPCS_ORACLE_OK;
}

/* Algorithm.HistogramByteComponents(ulong[] inArray, int l, int r)
 * Histogram.cs:957-990 -- eight counters per key; count is [8][256]. */
int hpcs_oracle_histogram_byte_components_u64(const uint64_t *a, int64_t l, int64_t r, int32_t *count)
{
    if (!a && r >= l) return HPCS_ORACLE_EINVAL;
    memset(count, 0, sizeof(int32_t) * 8 * 256);
    for (int64_t i = l; i <= r; i++) {
        uint64_t v = a[i];
        for (int d = 0; d < 8; d++)
            count[d * 256 + ((v >> (8 * d)) & 0xff)]++;
    }
    return HPCS_ORACLE_OK;
}

/* Algorithm.HistogramNbitComponents(uint[] inArray, int l, int r, int bits)
 * Histogram.cs:381-402 -- ceil(32/bits) digits of `bits` bits (the top digit
 * is narrower).  count is [D][2^bits]. */
int hpcs_oracle_histogram_nbit_components_u32(const uint32_t *a, int64_t l, int64_t r, int bits, int32_t *count)
{
    if (bits < 1 || bits > 16) return HPCS_ORACLE_EINVAL;
    int nbins = 1 << bits, ndig = (32 + bits - 1) / bits;
    uint32_t mask = (uint32_t)nbins - 1;
    memset(count, 0, sizeof(int32_t) * (size_t)ndig * nbins);
    for (int64_t i = l; i <= r; i++) {
        uint32_t v = a[i];
        for (int d = 0; d < ndig; d++)
            count[d * nbins + ((v >> (d * bits)) & mask)]++;
    }
    return HPCS_ORACLE_OK;
}

/* Algorithm.HistogramNbitComponents(ulong[] inArray, int l, int r, int bits)
 * Histogram.cs:404-427. */
int hpcs_oracle_histogram_nbit_components_u64(const uint64_t *a, int64_t l, int64_t r, int bits, int32_t *count)
{
    if (bits < 1 || bits > 16) return HPCS_ORACLE_EINVAL;
    int nbins = 1 << bits, ndig = (64 + bits - 1) / bits;
    uint64_t mask = (uint64_t)nbins - 1;
    memset(count, 0, sizeof(int32_t) * (size_t)ndig * nbins);
    for (int64_t i = l; i <= r; i++) {
        uint64_t v = a[i];
        for (int d = 0; d < ndig; d++)
            count[d * nbins + ((v >> (d * bits)) & mask)]++;
    }
    return HPCS_ORACLE_OK;
}

/* ParallelAlgorithm.HistogramByteComponentsParInner / Par (uint[])
 * HistogramParallel.cs:163-222: recursive halving at the overflow-free
 * midpoint m = r/2 + l/2 + (r%2 + l%2)/2, leaves of at most `threshold`
 * elements counted serially, children summed element-wise.  Parallel.Invoke
 * maps to an OpenMP task pair.  Counts are uint32 (the Par overloads return
 * uint[][]). */
static void hist_par_u32_rec(const uint32_t *a, int64_t l, int64_t r, int64_t thr, uint32_t *out)
{
    memset(out, 0, sizeof(uint32_t) * 4 * 256);
    if (l > r) return;
    if (r - l + 1 <= thr) {
        for (int64_t i = l; i <= r; i++) {
            uint32_t v = a[i];
            out[0 * 256 + (v & 0xff)]++;
            out[1 * 256 + ((v >> 8) & 0xff)]++;
            out[2 * 256 + ((v >> 16) & 0xff)]++;
            out[3 * 256 + ((v >> 24) & 0xff)]++;
        }
        return;
    }
    int64_t m = r / 2 + l / 2 + (r % 2 + l % 2) / 2;
    uint32_t right[4 * 256];
#pragma omp task shared(out) firstprivate(a, l, m, thr)
    hist_par_u32_rec(a, l, m, thr, out);
#pragma omp task shared(right) firstprivate(a, m, r, thr)
    hist_par_u32_rec(a, m + 1, r, thr, right);
#pragma omp taskwait
    for (int j = 0; j < 4 * 256; j++) out[j] += right[j];
}

static void hist_par_u64_rec(const uint64_t *a, int64_t l, int64_t r, int64_t thr, uint32_t *out)
{
    memset(out, 0, sizeof(uint32_t) * 8 * 256);
    if (l > r) return;
    if (r - l + 1 <= thr) {
        for (int64_t i = l; i <= r; i++) {
            uint64_t v = a[i];
            for (int d = 0; d < 8; d++) out[d * 256 + ((v >> (8 * d)) & 0xff)]++;
        }
        return;
    }
    int64_t m = r / 2 + l / 2 + (r % 2 + l % 2) / 2;
    uint32_t *right = (uint32_t *)malloc(sizeof(uint32_t) * 8 * 256);
#pragma omp task shared(out) firstprivate(a, l, m, thr)
    hist_par_u64_rec(a, l, m, thr, out);
#pragma omp task firstprivate(a, m, r, thr, right)
    hist_par_u64_rec(a, m + 1, r, thr, right);
#pragma omp taskwait
    for (int j = 0; j < 8 * 256; j++) out[j] += right[j];
    free(right);
}

/* HistogramParallel.cs:214-222: threshold = max(threshold, len / cores). */
int hpcs_oracle_histogram_byte_components_par_u32(const uint32_t *a, int64_t l, int64_t r, int64_t threshold, uint32_t *count)
{
    if (!a && r >= l) return HPCS_ORACLE_EINVAL;
    int64_t len = r - l + 1;
    int cores = omp_get_max_threads();
    if (threshold * cores < len) threshold = len / cores;
    if (threshold < 1) threshold = 1;
#pragma omp parallel
#pragma omp single
    hist_par_u32_rec(a, l, r, threshold, count);
    return HPCS_ORACLE_OK;
}

/* HistogramParallel.cs:487-555 (ulong[] overload). */
int hpcs_oracle_histogram_byte_components_par_u64(const uint64_t *a, int64_t l, int64_t r, int64_t threshold, uint32_t *count)
{
    if (!a && r >= l) return HPCS_ORACLE_EINVAL;
    int64_t len = r - l + 1;
    int cores = omp_get_max_threads();
    if (threshold * cores < len) threshold = len / cores;
    if (threshold < 1) threshold = 1;
#pragma omp parallel
#pragma omp single
    hist_par_u64_rec(a, l, r, threshold, count);
    return HPCS_ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Serial LSD radix sorts                                                    */
/* ------------------------------------------------------------------------ */

/* Algorithm.ComputeStartOfBins, RadixSortLsd.cs:112-125, with the base
 * offset of the Experimental overload (RadixSortLsd.cs:1436). */
static void start_of_bins(int ndig, int nbins, const int32_t *count, int64_t base, int64_t *start)
{
    for (int d = 0; d < ndig; d++) {
        start[(size_t)d * nbins] = base;
        for (int b = 1; b < nbins; b++)
            start[(size_t)d * nbins + b] = start[(size_t)d * nbins + b - 1] + count[(size_t)d * nbins + b - 1];
    }
}

/* Algorithm.SortRadixLsdInner(uint[], startIndex, length)
 * RadixSortLsd.cs:127-202: one histogram of all four byte digits, then four
 * stable scatter passes through 256 de-randomization buffers of 128 keys
 * (64 B cache line x 8 lines / 4 B), flushed whole when full (:165-186) and
 * partially after the pass (:187-194).  Ping-pong; four passes leave the
 * result in the caller's array. */
int hpcs_oracle_sort_radix_lsd_u32(uint32_t *a, int64_t start, int64_t len)
{
    if (!a && len > 0) return HPCS_ORACLE_EINVAL;
    if (len <= 1) return HPCS_ORACLE_OK;
    enum { BINS = 256, DEPTH = 64 * 8 / 4 };
    uint32_t *buf = (uint32_t *)malloc(sizeof(uint32_t) * BINS * DEPTH);
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)len);
    if (!buf || !tmp) { free(buf); free(tmp); return HPCS_ORACLE_ENOMEM; }
    int32_t count[4 * BINS];
    int64_t sob[4 * BINS];
    /* Work on the sub-array itself: offsets based at startIndex address the
     * same elements as offsets based at 0 address in a + startIndex. */
    a += start;
    int64_t end = len - 1;
    start = 0;
    hpcs_oracle_histogram_byte_components_u32(a, start, end, count);
    start_of_bins(4, BINS, count, start, sob);
    uint32_t *in = a, *out = tmp;
    int cur[BINS];
    for (int d = 0; d < 4; d++) {
        int shift = 8 * d;
        int64_t *s = sob + d * BINS;
        for (int b = 0; b < BINS; b++) cur[b] = b * DEPTH;
        for (int64_t i = start; i <= end; i++) {
            uint32_t v = in[i];
            uint32_t bin = (v >> shift) & 0xff;
            if (cur[bin] <= (int)(bin * DEPTH + DEPTH - 1)) {
                buf[cur[bin]++] = v;
            } else {
                memcpy(out + s[bin], buf + bin * DEPTH, sizeof(uint32_t) * DEPTH);
                s[bin] += DEPTH;
                cur[bin] = bin * DEPTH;
                buf[cur[bin]++] = v;
            }
        }
        for (int b = 0; b < BINS; b++) {
            int nitems = cur[b] - b * DEPTH;
            for (int k = 0; k < nitems; k++) out[s[b]++] = buf[b * DEPTH + k];
        }
        uint32_t *t = in; in = out; out = t;
    }
    free(buf);
    free(tmp);
    return HPCS_ORACLE_OK;
}

/* Algorithm.SortRadixLsdInner(ulong[], startIndex, length)
 * RadixSortLsd.cs:359-384: histogram of eight byte digits, eight plain
 * stable scatter passes; an even pass count leaves the result in place. */
int hpcs_oracle_sort_radix_lsd_u64(uint64_t *a, int64_t start, int64_t len)
{
    if (!a && len > 0) return HPCS_ORACLE_EINVAL;
    if (len <= 1) return HPCS_ORACLE_OK;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)len);
    if (!tmp) return HPCS_ORACLE_ENOMEM;
    int32_t count[8 * 256];
    int64_t sob[8 * 256];
    a += start;
    int64_t end = len - 1;
    start = 0;
    hpcs_oracle_histogram_byte_components_u64(a, start, end, count);
    start_of_bins(8, 256, count, start, sob);
    uint64_t *in = a, *out = tmp;
    for (int d = 0; d < 8; d++) {
        int64_t *s = sob + d * 256;
        for (int64_t i = start; i <= end; i++)
            out[s[(in[i] >> (8 * d)) & 0xff]++] = in[i];
        uint64_t *t = in; in = out; out = t;
    }
    free(tmp);
    return HPCS_ORACLE_OK;
}

/* HPCsharpExperimental.Algorithm.SortRadixLsdNbit(uint[], start, length, bits)
 * RadixSortLsd.cs:1616-1661: N-bit digits from HistogramNbitComponents, plain
 * stable passes; an odd pass count copies the result back (:1659-1660). */
int hpcs_oracle_sort_radix_lsd_nbit_u32(uint32_t *a, int64_t start, int64_t len, int bits)
{
    if (!a && len > 0) return HPCS_ORACLE_EINVAL;
    if (bits < 1 || bits > 16) return HPCS_ORACLE_EINVAL;
    if (len <= 1) return HPCS_ORACLE_OK;
    int nbins = 1 << bits, ndig = (32 + bits - 1) / bits;
    uint32_t mask = (uint32_t)nbins - 1;
    uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)len);
    int32_t *count = (int32_t *)malloc(sizeof(int32_t) * (size_t)ndig * nbins);
    int64_t *sob = (int64_t *)malloc(sizeof(int64_t) * (size_t)ndig * nbins);
    if (!tmp || !count || !sob) { free(tmp); free(count); free(sob); return HPCS_ORACLE_ENOMEM; }
    a += start;
    int64_t end = len - 1;
    start = 0;
    hpcs_oracle_histogram_nbit_components_u32(a, start, end, bits, count);
    start_of_bins(ndig, nbins, count, start, sob);
    uint32_t *in = a, *out = tmp;
    for (int d = 0; d < ndig; d++) {
        int64_t *s = sob + (size_t)d * nbins;
        for (int64_t i = start; i <= end; i++)
            out[s[(in[i] >> (d * bits)) & mask]++] = in[i];
        uint32_t *t = in; in = out; out = t;
    }
    if (ndig & 1) memcpy(a + start, in + start, sizeof(uint32_t) * (size_t)len);
    free(tmp); free(count); free(sob);
    return HPCS_ORACLE_OK;
}

/* N-bit LSD for ulong keys: the ulong counterpart of SortRadixLsdNbit
 * (RadixSortLsd.cs:1616-1661 over HistogramNbitComponents(ulong[]),
 * Histogram.cs:404-427). */
int hpcs_oracle_sort_radix_lsd_nbit_u64(uint64_t *a, int64_t start, int64_t len, int bits)
{
    if (!a && len > 0) return HPCS_ORACLE_EINVAL;
    if (bits < 1 || bits > 16) return HPCS_ORACLE_EINVAL;
    if (len <= 1) return HPCS_ORACLE_OK;
    int nbins = 1 << bits, ndig = (64 + bits - 1) / bits;
    uint64_t mask = (uint64_t)nbins - 1;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)len);
    int32_t *count = (int32_t *)malloc(sizeof(int32_t) * (size_t)ndig * nbins);
    int64_t *sob = (int64_t *)malloc(sizeof(int64_t) * (size_t)ndig * nbins);
    if (!tmp || !count || !sob) { free(tmp); free(count); free(sob); return HPCS_ORACLE_ENOMEM; }
    a += start;
    int64_t end = len - 1;
    start = 0;
    hpcs_oracle_histogram_nbit_components_u64(a, start, end, bits, count);
    start_of_bins(ndig, nbins, count, start, sob);
    uint64_t *in = a, *out = tmp;
    for (int d = 0; d < ndig; d++) {
        int64_t *s = sob + (size_t)d * nbins;
        for (int64_t i = start; i <= end; i++)
            out[s[(in[i] >> (d * bits)) & mask]++] = in[i];
        uint64_t *t = in; in = out; out = t;
    }
    if (ndig & 1) memcpy(a + start, in + start, sizeof(uint64_t) * (size_t)len);
    free(tmp); free(count); free(sob);
    return HPCS_ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Parallel LSD radix sort (the CPU baseline)                                */
/* ------------------------------------------------------------------------ */

/* Algorithm.HistogramByteComponentsAcrossWorkQuantasQC, Histogram.cs:574-637:
 * per work quantum, a 256-bin count of one byte digit over [l, r]. */
static void hist_qc_u32(const uint32_t *a, int64_t l, int64_t r, int64_t wq, int whichByte, uint32_t *count /* [Q][256] */)
{
    int shift = 8 * whichByte;
    for (int64_t i = l; i <= r; i++)
        count[(i / wq) * 256 + ((a[i] >> shift) & 0xff)]++;
}

/* ParallelAlgorithm.HistogramByteComponentsQCParInner, HistogramParallel.cs:
 * 285-319: halving with leaf threshold max(16Ki, n/cores); each leaf counts
 * into its own [Q][256] array, the reduce adds only the quanta the two halves
 * touched (:303-307). */
static void hist_qc_par_rec(const uint32_t *a, int64_t l, int64_t r, int64_t wq, int64_t nq, int whichByte, int64_t thr, uint32_t *out)
{
    if (r - l + 1 <= thr) {
        hist_qc_u32(a, l, r, wq, whichByte, out);
        return;
    }
    int64_t m = r / 2 + l / 2 + (r % 2 + l % 2) / 2;
    uint32_t *right = (uint32_t *)calloc((size_t)nq * 256, sizeof(uint32_t));
#pragma omp task shared(out) firstprivate(a, l, m, wq, nq, whichByte, thr)
    hist_qc_par_rec(a, l, m, wq, nq, whichByte, thr, out);
#pragma omp task firstprivate(a, m, r, wq, nq, whichByte, thr, right)
    hist_qc_par_rec(a, m + 1, r, wq, nq, whichByte, thr, right);
#pragma omp taskwait
    for (int64_t q = l / wq; q <= r / wq; q++)
        for (int j = 0; j < 256; j++) out[q * 256 + j] += right[q * 256 + j];
    free(right);
}

/* ParallelAlgorithm.SortRadixPar(uint[] inputArray, int ParallelWorkQuantum = 64Ki)
 * RadixSortLsdParallel.cs:145-232.  Per digit pass:
 *   ComputeStartOfBinsPar (:49-94): per-quantum counts (HistogramByteComponentsQCPar),
 *     bin-major / quantum-minor offsets start[q][b] = start[q-1][b] + count[q-1][b];
 *   one task per quantum (:186-215) -> SortRadixDeRandomizedInnerFunction (:96-135):
 *     stable scatter through 64-deep per-bin buffers, flushed when full and at
 *     the end of the quantum;
 *   Task.WaitAll pass barrier (:216), then swap.
 * Four passes leave the result in the caller's array (:223-229).
 * The buffers and scratch are allocated per call, as in the reference. */
int hpcs_oracle_sort_radix_par_u32(uint32_t *a, int64_t n, int64_t quantum)
{
    if (!a && n > 0) return HPCS_ORACLE_EINVAL;
    if (quantum <= 0) return HPCS_ORACLE_EINVAL;
    if (n == 0) return HPCS_ORACLE_OK;
    enum { BINS = 256, DEPTH = 64 };
    int64_t nq = (n + quantum - 1) / quantum;
    uint32_t *out = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t *bufs = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nq * BINS * DEPTH);
    uint32_t *count = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nq * BINS);
    uint32_t *sob = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)nq * BINS);
    if (!out || !bufs || !count || !sob) { free(out); free(bufs); free(count); free(sob); return HPCS_ORACLE_ENOMEM; }
    int cores = omp_get_max_threads();
    int64_t thr = 16 * 1024;
    if (thr * cores < n) thr = n / cores;
    uint32_t *in = a, *dst = out;
    for (int digit = 0; digit < 4; digit++) {
        memset(count, 0, sizeof(uint32_t) * (size_t)nq * BINS);
#pragma omp parallel
#pragma omp single
        hist_qc_par_rec(in, 0, n - 1, quantum, nq, digit, thr, count);
        uint32_t size[BINS];
        for (int b = 0; b < BINS; b++) {
            uint32_t s = 0;
            for (int64_t q = 0; q < nq; q++) s += count[q * BINS + b];
            size[b] = s;
        }
        sob[0] = 0;
        for (int b = 1; b < BINS; b++) sob[b] = sob[b - 1] + size[b - 1];
        for (int64_t q = 1; q < nq; q++)
            for (int b = 0; b < BINS; b++) sob[q * BINS + b] = sob[(q - 1) * BINS + b] + count[(q - 1) * BINS + b];
        int shift = 8 * digit;
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t q = 0; q < nq; q++) {
            uint32_t *s = sob + q * BINS;
            uint32_t *buf = bufs + q * BINS * DEPTH;
            uint32_t idx[BINS];
            for (int b = 0; b < BINS; b++) idx[b] = (uint32_t)b * DEPTH;
            int64_t lo = q * quantum, hi = lo + quantum < n ? lo + quantum : n;
            for (int64_t i = lo; i < hi; i++) {
                uint32_t v = in[i];
                uint32_t dg = (v >> shift) & 0xff;
                if (idx[dg] < (dg + 1) * DEPTH) {
                    buf[idx[dg]++] = v;
                } else {
                    memcpy(dst + s[dg], buf + dg * DEPTH, sizeof(uint32_t) * DEPTH);
                    s[dg] += DEPTH;
                    buf[dg * DEPTH] = v;
                    idx[dg] = dg * DEPTH + 1;
                }
            }
            for (int b = 0; b < BINS; b++) {
                uint32_t bs = (uint32_t)b * DEPTH;
                while (bs < idx[b]) dst[s[b]++] = buf[bs++];
            }
        }
        uint32_t *t = in; in = dst; dst = t;
    }
    free(out); free(bufs); free(count); free(sob);
    return HPCS_ORACLE_OK;
}

/* ParallelAlgorithm.SortRadixPar(ulong[]) RadixSortLsdParallel.cs:690-734:
 * parallel histogram (HistogramByteComponentsPar, HistogramParallel.cs:
 * 487-555), serial stable scatter, eight passes. */
int hpcs_oracle_sort_radix_par_u64(uint64_t *a, int64_t n)
{
    if (!a && n > 0) return HPCS_ORACLE_EINVAL;
    if (n <= 1) return HPCS_ORACLE_OK;
    uint64_t *tmp = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    if (!tmp) return HPCS_ORACLE_ENOMEM;
    uint32_t count[8 * 256];
    uint32_t sob[8 * 256];
    hpcs_oracle_histogram_byte_components_par_u64(a, 0, n - 1, 16 * 1024, count);
    for (int d = 0; d < 8; d++) {
        sob[d * 256] = 0;
        for (int b = 1; b < 256; b++) sob[d * 256 + b] = sob[d * 256 + b - 1] + count[d * 256 + b - 1];
    }
    uint64_t *in = a, *out = tmp;
    for (int d = 0; d < 8; d++) {
        uint32_t *s = sob + d * 256;
        for (int64_t i = 0; i < n; i++) out[s[(in[i] >> (8 * d)) & 0xff]++] = in[i];
        uint64_t *t = in; in = out; out = t;
    }
    free(tmp);
    return HPCS_ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Floating-point keys                                                       */
/* ------------------------------------------------------------------------ */

/* Radix key of a float's bits, RadixSortMsd.cs:1379-1383 (and the per-digit
 * form :1405-1409, Histogram.cs:1224-1241): non-negative values flip the sign
 * bit, negative values flip every bit.  Ascending keys give ascending values
 * with -0.0 before +0.0. */
uint32_t hpcs_oracle_f32_key(uint32_t bits)
{
    return (bits & 0x80000000u) ? ~bits : (bits ^ 0x80000000u);
}

uint32_t hpcs_oracle_f32_unkey(uint32_t key)
{
    return (key & 0x80000000u) ? (key ^ 0x80000000u) : ~key;
}

/* Double form, RadixSortMsd.cs:1484-1488, :1510-1514; Histogram.cs:1264-1281. */
uint64_t hpcs_oracle_f64_key(uint64_t bits)
{
    return (bits & 0x8000000000000000ull) ? ~bits : (bits ^ 0x8000000000000000ull);
}

uint64_t hpcs_oracle_f64_unkey(uint64_t key)
{
    return (key & 0x8000000000000000ull) ? (key ^ 0x8000000000000000ull) : ~key;
}

/* SortRadixMsd(float[]) RadixSortMsd.cs:1563-1571 orders floats by the key
 * above.  The key is a bijection on bit patterns, so (for NaN-free input,
 * where the reference's Array.Sort base case agrees with the key order) the
 * reference's output equals the LSD sort of the keys, mapped back. */
int hpcs_oracle_sort_radix_f32(float *a, int64_t n)
{
    if (!a && n > 0) return HPCS_ORACLE_EINVAL;
    uint32_t *u = (uint32_t *)a;
    for (int64_t i = 0; i < n; i++) u[i] = hpcs_oracle_f32_key(u[i]);
    int rc = hpcs_oracle_sort_radix_lsd_u32(u, 0, n);
    for (int64_t i = 0; i < n; i++) u[i] = hpcs_oracle_f32_unkey(u[i]);
    return rc;
}

/* SortRadixMsd(double[]) RadixSortMsd.cs:1587-1595, same reasoning. */
int hpcs_oracle_sort_radix_f64(double *a, int64_t n)
{
    if (!a && n > 0) return HPCS_ORACLE_EINVAL;
    uint64_t *u = (uint64_t *)a;
    for (int64_t i = 0; i < n; i++) u[i] = hpcs_oracle_f64_key(u[i]);
    int rc = hpcs_oracle_sort_radix_lsd_u64(u, 0, n);
    for (int64_t i = 0; i < n; i++) u[i] = hpcs_oracle_f64_unkey(u[i]);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Records sorted by an extracted key (key-extractor overloads)              */
/* ------------------------------------------------------------------------ */

#define HPCS_KEY_UINT 0
#define HPCS_KEY_FLOAT 1
#define HPCS_KEY_INT 2

static uint64_t record_key(const unsigned char *rec, int key_offset, int key_bits, int key_kind)
{
    if (key_bits == 32) {
        uint32_t k;
        memcpy(&k, rec + key_offset, 4);
        return key_kind == HPCS_KEY_FLOAT ? hpcs_oracle_f32_key(k) : key_kind == HPCS_KEY_INT ? (k ^ 0x80000000u) : k;
    } else {
        uint64_t k;
        memcpy(&k, rec + key_offset, 8);
        return key_kind == HPCS_KEY_FLOAT ? hpcs_oracle_f64_key(k) : key_kind == HPCS_KEY_INT ? (k ^ 0x8000000000000000ull) : k;
    }
}

/* Algorithm.SortRadixLsd2<T>(T[], Func<T,UInt32>) RadixSortLsd.cs:751-775 and
 * SortRadixLsd2<T>(T[], Func<T,UInt64>) RadixSortLsd.cs:821-846: one byte
 * histogram of the extracted key (Histogram.cs:760-844), then 4 or 8 stable
 * passes that move whole records, each pass extracting the key again.  The
 * record is opaque bytes of rec_size; the key sits at key_offset. */
int hpcs_oracle_sort_radix_records(void *recs, int64_t n, int32_t rec_size, int32_t key_offset, int32_t key_bits, int32_t key_kind)
{
    if (!recs && n > 0) return HPCS_ORACLE_EINVAL;
    if (key_bits != 32 && key_bits != 64) return HPCS_ORACLE_EINVAL;
    if (key_offset < 0 || key_offset + key_bits / 8 > rec_size) return HPCS_ORACLE_EINVAL;
    if (key_kind != HPCS_KEY_UINT && key_kind != HPCS_KEY_FLOAT && key_kind != HPCS_KEY_INT) return HPCS_ORACLE_EINVAL;
    if (n <= 1) return HPCS_ORACLE_OK;
    int ndig = key_bits / 8;
    unsigned char *tmp = (unsigned char *)malloc((size_t)rec_size * (size_t)n);
    if (!tmp) return HPCS_ORACLE_ENOMEM;
    int32_t count[8 * 256];
    int64_t sob[8 * 256];
    memset(count, 0, sizeof(count));
    unsigned char *in = (unsigned char *)recs, *out = tmp;
    for (int64_t i = 0; i < n; i++) {
        uint64_t k = record_key(in + (size_t)i * rec_size, key_offset, key_bits, key_kind);
        for (int d = 0; d < ndig; d++) count[d * 256 + ((k >> (8 * d)) & 0xff)]++;
    }
    start_of_bins(ndig, 256, count, 0, sob);
    for (int d = 0; d < ndig; d++) {
        int64_t *s = sob + d * 256;
        for (int64_t i = 0; i < n; i++) {
            uint64_t k = record_key(in + (size_t)i * rec_size, key_offset, key_bits, key_kind);
            memcpy(out + (size_t)(s[(k >> (8 * d)) & 0xff]++) * rec_size, in + (size_t)i * rec_size, (size_t)rec_size);
        }
        unsigned char *t = in; in = out; out = t;
    }
    free(tmp);
    return HPCS_ORACLE_OK;
}

/* Key array plus items, Algorithm.SortRadix<T>(T[] items, UInt64[] keys)
 * RadixSortLsd.cs:1245-1299 restated with the full 64-bit key (the
 * reference's 32-bit mask loop stops after 4 passes, :1258,1284): keys and
 * the uint32 items move together through stable passes. */
int hpcs_oracle_sort_radix_pairs_u64_u32(uint64_t *keys, uint32_t *vals, int64_t n)
{
    if ((!keys || !vals) && n > 0) return HPCS_ORACLE_EINVAL;
    if (n <= 1) return HPCS_ORACLE_OK;
    uint64_t *tk = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    if (!tk || !tv) { free(tk); free(tv); return HPCS_ORACLE_ENOMEM; }
    int32_t count[8 * 256];
    int64_t sob[8 * 256];
    hpcs_oracle_histogram_byte_components_u64(keys, 0, n - 1, count);
    start_of_bins(8, 256, count, 0, sob);
    uint64_t *ki = keys, *ko = tk;
    uint32_t *vi = vals, *vo = tv;
    for (int d = 0; d < 8; d++) {
        int64_t *s = sob + d * 256;
        for (int64_t i = 0; i < n; i++) {
            int64_t p = s[(ki[i] >> (8 * d)) & 0xff]++;
            ko[p] = ki[i];
            vo[p] = vi[i];
        }
        uint64_t *t = ki; ki = ko; ko = t;
        uint32_t *tt = vi; vi = vo; vo = tt;
    }
    free(tk); free(tv);
    return HPCS_ORACLE_OK;
}

/* Algorithm.SortRadix<T>(T[] items, UInt32[] keys) RadixSortLsd.cs:1183-1237:
 * uint32 keys with uint32 items, four stable passes. */
int hpcs_oracle_sort_radix_pairs_u32_u32(uint32_t *keys, uint32_t *vals, int64_t n)
{
    if ((!keys || !vals) && n > 0) return HPCS_ORACLE_EINVAL;
    if (n <= 1) return HPCS_ORACLE_OK;
    uint32_t *tk = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    uint32_t *tv = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)n);
    if (!tk || !tv) { free(tk); free(tv); return HPCS_ORACLE_ENOMEM; }
    int32_t count[4 * 256];
    int64_t sob[4 * 256];
    hpcs_oracle_histogram_byte_components_u32(keys, 0, n - 1, count);
    start_of_bins(4, 256, count, 0, sob);
    uint32_t *ki = keys, *ko = tk, *vi = vals, *vo = tv;
    for (int d = 0; d < 4; d++) {
        int64_t *s = sob + d * 256;
        for (int64_t i = 0; i < n; i++) {
            int64_t p = s[(ki[i] >> (8 * d)) & 0xff]++;
            ko[p] = ki[i];
            vo[p] = vi[i];
        }
        uint32_t *t = ki; ki = ko; ko = t;
        t = vi; vi = vo; vo = t;
    }
    free(tk); free(tv);
    return HPCS_ORACLE_OK;
}

/* ------------------------------------------------------------------------ */
/* Sub-range (segment) sorts                                                 */
/* ------------------------------------------------------------------------ */

/* HPCsharpExperimental.Algorithm.SortRadixLsd(uint[], startIndex, length)
 * RadixSortLsd.cs:1409-1457 sorts a[start, start+len) in place with offsets
 * based at startIndex (:1436).  A batch of segments is that call per segment;
 * segment s is [seg_begin[s], seg_begin[s+1]) and seg_begin has nseg+1
 * entries; keys outside every segment are untouched. */
int hpcs_oracle_sort_segments_u32(uint32_t *a, const int64_t *seg_begin, int32_t nseg)
{
    if (nseg < 0 || (nseg > 0 && (!a || !seg_begin))) return HPCS_ORACLE_EINVAL;
    for (int32_t s = 0; s < nseg; s++)
        if (seg_begin[s + 1] < seg_begin[s]) return HPCS_ORACLE_EINVAL;
    int rc = HPCS_ORACLE_OK;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc)
    for (int32_t s = 0; s < nseg; s++) {
        int r = hpcs_oracle_sort_radix_lsd_u32(a, seg_begin[s], seg_begin[s + 1] - seg_begin[s]);
        if (r < rc) rc = r;
    }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Radix top-k selection                                                     */
/* ------------------------------------------------------------------------ */

/* Algorithm.MoveRobElementsIntoKthBin, SelectionTopKRadix.cs:19-41: move the
 * elements right of the k-th bin that belong in it into it, swapping out
 * elements that belong to the right. */
static int64_t topk_move_rob_into(uint32_t *a, int64_t sob_k, int64_t len_k, int64_t s_rob, int64_t len_rob, int shift,
                                  uint32_t mask, uint32_t kb)
{
    int64_t end_k = sob_k + len_k - 1, end_rob = s_rob + len_rob - 1;
    int64_t cr = s_rob, ci = sob_k;
    for (;;) {
        for (; cr <= end_rob; cr++)
            if (((a[cr] >> shift) & mask) == kb) break;
        if (cr <= end_rob)
            for (; ci <= end_k; ci++)
                if (((a[ci] >> shift) & mask) != kb) break;
        if (cr > end_rob || ci > end_k) break;
        if (((a[ci] >> shift) & mask) > kb) {
            uint32_t t = a[ci];
            a[ci++] = a[cr];
            a[cr++] = t;
        } else {
            a[ci++] = a[cr++];
        }
    }
    return ci;
}

/* Algorithm.MoveRobElementsOutOfKthBin, SelectionTopKRadix.cs:44-60. */
static void topk_move_rob_out(uint32_t *a, int64_t sob_k, int64_t len_k, int64_t s_rob, int64_t len_rob, int shift,
                              uint32_t mask, uint32_t kb)
{
    int64_t end_k = sob_k + len_k - 1, end_rob = s_rob + len_rob - 1;
    int64_t cr = s_rob, ci = sob_k;
    for (;;) {
        for (; ci <= end_k; ci++)
            if (((a[ci] >> shift) & mask) > kb) break;
        if (ci <= end_k)
            for (; cr <= end_rob; cr++)
                if (((a[cr] >> shift) & mask) < kb) break;
        if (cr > end_rob || ci > end_k) break;
        a[cr++] = a[ci++];
    }
}

/* Algorithm.MoveLobElements, SelectionTopKRadix.cs:64-93. */
static void topk_move_lob(uint32_t *a, int64_t s_lob, int64_t len_lob, int64_t sob_k, int64_t len_k, int64_t s_kp1,
                          int64_t len_rob, int shift, uint32_t mask, uint32_t kb)
{
    int64_t end_k = sob_k + len_k - 1, end_lob = s_lob + len_lob - 1, end_rob = s_kp1 + len_rob - 1;
    int64_t cl = s_lob, ci = sob_k, cr = s_kp1;
    for (;;) {
        int found_ib = 0, found_rob = 0;
        for (; cl <= end_lob; cl++) {
            if (((a[cl] >> shift) & mask) == kb) { found_ib = 1; break; }
            if (((a[cl] >> shift) & mask) > kb) { found_rob = 1; break; }
        }
        if (cl > end_lob) break;
        if (found_ib) {
            for (; ci <= end_k; ci++)
                if (((a[ci] >> shift) & mask) != kb) break;
            a[ci++] = a[cl++];
        }
        if (found_rob) {
            for (; cr <= end_rob; cr++)
                if (((a[cr] >> shift) & mask) <= kb) break;
            a[cr++] = a[cl++];
        }
    }
}

/* Algorithm.RadixSelectionTopKWordInner, SelectionTopKRadix.cs:96-133:
 * 16-bit MSD digits; the bin holding index k is gathered in place, elements
 * of higher bins moved to its right, and the recursion continues in that bin
 * with the next 16 bits. */
static int topk_word_inner(uint32_t *a, int64_t first, int64_t length, int shift, int64_t k, int32_t *count,
                           int64_t *sob)
{
    const int BINS = 65536;
    const uint32_t mask = 0xFFFF;
    int64_t last = first + length - 1;
    memset(count, 0, sizeof(int32_t) * BINS);
    for (int64_t i = first; i <= last; i++) count[(uint16_t)(a[i] >> shift)]++;
    sob[0] = first;
    sob[BINS] = last + 1;
    for (int i = 1; i < BINS; i++) sob[i] = sob[i - 1] + count[i - 1];
    int kb = 0;
    for (; kb < BINS; kb++) {
        int64_t bl = sob[kb + 1] - sob[kb];
        if (bl == 0) continue;
        if (k >= sob[kb] && k <= sob[kb + 1] - 1) break;
    }
    if (kb >= BINS) return HPCS_ORACLE_EINVAL;
    int64_t len_k = sob[kb + 1] - sob[kb], s_kp1 = sob[kb + 1], s_k = sob[kb];
    int64_t ci = topk_move_rob_into(a, s_k, len_k, s_kp1, last - s_kp1 + 1, shift, mask, (uint32_t)kb);
    topk_move_rob_out(a, ci, s_kp1 - ci, s_kp1, last - s_kp1 + 1, shift, mask, (uint32_t)kb);
    topk_move_lob(a, first, s_k - first, s_k, len_k, s_kp1, last - s_kp1 + 1, shift, mask, (uint32_t)kb);
    if (shift > 0) {
        shift = shift >= 16 ? shift - 16 : 0;
        if (len_k > 1) return topk_word_inner(a, s_k, len_k, shift, k, count, sob);
        if (len_k == 1) return HPCS_ORACLE_OK;
        return HPCS_ORACLE_EINVAL; /* the reference throws here (:129) */
    }
    return HPCS_ORACLE_OK;
}

/* Algorithm.SelectTopK(this uint[] a, int k), SelectionTopKRadix.cs:160-171:
 * afterwards a[k] is the k-th smallest key (0-based) and a[k..n) holds the
 * n-k largest keys.  Elements are moved rather than swapped, so a[0..k) is
 * left unspecified, as in the reference. */
int hpcs_oracle_select_topk_u32(uint32_t *a, int64_t n, int64_t k)
{
    if (!a || n <= 0 || k < 0 || k >= n) return HPCS_ORACLE_EINVAL;
    int32_t *count = (int32_t *)malloc(sizeof(int32_t) * 65536);
    int64_t *sob = (int64_t *)malloc(sizeof(int64_t) * 65537);
    if (!count || !sob) { free(count); free(sob); return HPCS_ORACLE_ENOMEM; }
    int rc = topk_word_inner(a, 0, n, 16, k, count, sob);
    free(count);
    free(sob);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Signed keys                                                               */
/* ------------------------------------------------------------------------ */

/* Algorithm.HistogramByteComponents(int[], l, r), Histogram.cs:992-1017: the
 * low three bytes as they are, the top byte ^ 128. */
int hpcs_oracle_histogram_byte_components_i32(const int32_t *a, int64_t l, int64_t r, int32_t *count)
{
    if (!a || !count) return HPCS_ORACLE_EINVAL;
    memset(count, 0, sizeof(int32_t) * 4 * 256);
    for (int64_t i = l; i <= r; i++) {
        uint32_t u = (uint32_t)a[i];
        count[0 * 256 + (u & 0xff)]++;
        count[1 * 256 + ((u >> 8) & 0xff)]++;
        count[2 * 256 + ((u >> 16) & 0xff)]++;
        count[3 * 256 + ((u >> 24) ^ 128)]++;
    }
    return HPCS_ORACLE_OK;
}

/* HistogramByteComponents(long[], l, r), Histogram.cs:1019-1052. */
int hpcs_oracle_histogram_byte_components_i64(const int64_t *a, int64_t l, int64_t r, int32_t *count)
{
    if (!a || !count) return HPCS_ORACLE_EINVAL;
    memset(count, 0, sizeof(int32_t) * 8 * 256);
    for (int64_t i = l; i <= r; i++) {
        uint64_t u = (uint64_t)a[i];
        for (int d = 0; d < 7; d++) count[d * 256 + ((u >> (8 * d)) & 0xff)]++;
        count[7 * 256 + ((u >> 56) ^ 128)]++;
    }
    return HPCS_ORACLE_OK;
}

/* SortRadixLsdInner(this int[], start, length), RadixSortLsd.cs:267-300 (the
 * long[] form :444-532): four (eight) stable byte passes whose last digit is
 * the top byte ^ 128; the result is back in a after an even pass count. */
int hpcs_oracle_sort_radix_lsd_i32(int32_t *a, int64_t n)
{
    if (!a && n > 0) return HPCS_ORACLE_EINVAL;
    if (n <= 1) return HPCS_ORACLE_OK;
    int32_t count[4 * 256];
    int64_t sob[4 * 256];
    hpcs_oracle_histogram_byte_components_i32(a, 0, n - 1, count);
    start_of_bins(4, 256, count, 0, sob);
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    if (!tmp) return HPCS_ORACLE_ENOMEM;
    int32_t *in = a, *out = tmp;
    for (int d = 0; d < 4; d++) {
        int64_t *s = sob + d * 256;
        for (int64_t i = 0; i < n; i++) {
            uint32_t u = (uint32_t)in[i];
            uint32_t dig = d < 3 ? ((u >> (8 * d)) & 0xff) : ((u >> 24) ^ 128);
            out[s[dig]++] = in[i];
        }
        int32_t *t = in; in = out; out = t;
    }
    free(tmp);
    return HPCS_ORACLE_OK;
}

int hpcs_oracle_sort_radix_lsd_i64(int64_t *a, int64_t n)
{
    if (!a && n > 0) return HPCS_ORACLE_EINVAL;
    if (n <= 1) return HPCS_ORACLE_OK;
    int32_t count[8 * 256];
    int64_t sob[8 * 256];
    hpcs_oracle_histogram_byte_components_i64(a, 0, n - 1, count);
    start_of_bins(8, 256, count, 0, sob);
    int64_t *tmp = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    if (!tmp) return HPCS_ORACLE_ENOMEM;
    int64_t *in = a, *out = tmp;
    for (int d = 0; d < 8; d++) {
        int64_t *s = sob + d * 256;
        for (int64_t i = 0; i < n; i++) {
            uint64_t u = (uint64_t)in[i];
            uint64_t dig = d < 7 ? ((u >> (8 * d)) & 0xff) : ((u >> 56) ^ 128);
            out[s[dig]++] = in[i];
        }
        int64_t *t = in; in = out; out = t;
    }
    free(tmp);
    return HPCS_ORACLE_OK;
}
