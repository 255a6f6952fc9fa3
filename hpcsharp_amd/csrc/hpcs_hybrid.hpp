// hpcs_hybrid.hpp -- the MSD-first form of the 4-byte unsigned keys-only
// sort (pass algorithm 6).  Same result as the LSD passes of
// SortRadixLsd(uint[]) (RadixSortLsd.cs:127-260): the keys in ascending
// order; keys alone carry no identity, so the order in which equal keys are
// placed does not show and the stages below need not be stable.
//
// Sequence (every launch planned on the device, graph-capturable):
//   k_upsweep<TOP>   one read: OR / AND of the keys, per-tile counts of byte 0
//                    and of byte 3
//   k_plan_rts       if a bit from 23 up varies: the MSD-first form, split
//                    digits s1 = min(24, top varying bit - 7) and s2 = s1 - 8
//                    (HYB_MSD); otherwise the LSD passes over the varying
//                    bytes, one per stage slot below (HYB_LSD)
//   global pass      pass slot 3 of the plan: tile counts (the upsweep's when
//                    the digit is the top byte or byte 0), range sums, column
//                    scan, scatter A -> B by the 8 bits at s1 (256 buckets)
//   k_hyb_begin1     the segments of the segmented pass: the buckets
//   level 2          the batched segments' machinery (hpcs_segments.hpp) on
//                    the 256 buckets in B: buckets of at most SEG_SMALL keys
//                    sorted whole in LDS into A, larger ones split by the 8
//                    bits at s2 (counts, scan, scatter B -> A) into 256
//                    sub-buckets each
//   k_hyb_local      one workgroup per sub-bucket, in place in A: its keys
//                    share bits 16..31, so a counting pass over the 12-bit
//                    bins below s2 in LDS and a rank within each bin (at most
//                    HYB_MAX_BIN keys) finish it: one read and one write per
//                    key
//   overflow         sub-buckets beyond HYB_CAP keys or with a larger bin:
//                    up to SEG_MEDIUM keys one 1024-thread workgroup in LDS,
//                    larger ones two segmented passes (bits 0-7, 8-15: A -> B
//                    -> A)
// With HYB_LSD the lowest varying byte takes the global pass, the second the
// segmented pass over one segment of all keys, the third the first overflow
// pass, and a copy back follows an odd number of them; each slot without a
// byte does nothing.
// Bytes moved per key: upsweep 4, top-byte pass 8, byte-2 pass 12 (count,
// read, write), local sort 8 = 32 (the LSD passes move 48 in this form: one
// counting read per later pass).
#pragma once
#include "hpcs_segments.hpp"

namespace hpcs {

constexpr int HYB_BLOCK = 256;
constexpr int HYB_ITEMS = 20;
constexpr int HYB_CAP = HYB_BLOCK * HYB_ITEMS; // largest sub-bucket of the local sort
constexpr int HYB_NBIN = 4096;                 // 12-bit bins: key bits 4..15
constexpr int HYB_MAX_BIN = 256;               // keys per bin the rank loop takes
constexpr int HYB_SUB = RADIX * RADIX;         // sub-buckets: one per (byte 3, byte 2)
// Batched segments: segments longer than this many keys take the MSD-first
// form by default (its 256 buckets then hold at least 256 keys on average)
constexpr int64_t SEG_MSD_DEFAULT = 65536;
constexpr int SEGMSD_BLOCK = 256; // the bucket kernel's workgroup: 256 x 20 keys
constexpr int SEGMSD_ITEMS = 20;

// The segments of the segmented pass (begin1[0..256]) and a clear overflow
// plan (sp2).  MSD-first form: the 256 buckets of the first split, from pass
// 3's digit starts.  LSD passes with a second varying byte: segment 0 holds
// every key.  Otherwise every segment is empty, so the later stages do
// nothing.  One workgroup of RADIX threads.
static __global__ __launch_bounds__(RADIX) void k_hyb_begin1(const DevPlan *__restrict__ plan, uint64_t n,
                                                          int64_t *__restrict__ begin1, SegPlan *__restrict__ sp2)
{
    static_assert(sizeof(SegPlan) / 4 <= RADIX, "one thread per word of the overflow plan");
    const int d = threadIdx.x;
    const uint32_t mode = plan->hybrid;
    const bool whole = mode == HYB_LSD && plan->hyb_shift2 < HYB_SKIP;
    begin1[d] = mode == HYB_MSD ? int64_t(plan->gstart[3][d]) : (whole && d > 0) ? int64_t(n) : int64_t(0);
    if (d == 0) begin1[RADIX] = (mode == HYB_MSD || whole) ? int64_t(n) : int64_t(0);
    if (d < int(sizeof(SegPlan) / 4)) reinterpret_cast<uint32_t *>(sp2)[d] = 0u;
}

// LDS slot of bin b: one padding word per 16 bins, so the scan's thread t
// (bins 16t .. 16t + 15) meets no bank conflict at any step.
__device__ __forceinline__ int hyb_slot(int b) { return b + (b >> 4); }


// Workgroup b = k * RADIX + d sorts sub-bucket d (byte 2 = d) of the k-th
// large bucket of level 2 (sp, large_list: level 2's plan; dstart[k][d]: the
// digit starts within the bucket from level 2's scan, k_segx_bases).
// Sub-buckets of at most CAP keys whose 12-bit bins hold at
// most HYB_MAX_BIN keys are sorted here, in place; the others are listed in
// the overflow plan sp2 as (begin, end) pairs: pair i at pairs[2i], listed
// as segment 2i, medium (at most SEG_MEDIUM keys) or large (with its first
// tile in tile0_2, tiles of `tile` keys).
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_hyb_local(uint32_t *__restrict__ keys, const int64_t *__restrict__ begin1,
                                                     const DevPlan *__restrict__ plan, const SegPlan *__restrict__ sp,
                                                     const uint32_t *__restrict__ large_list,
                                                     const uint32_t *__restrict__ dstart, SegPlan *__restrict__ sp2,
                                                     int64_t *__restrict__ pairs, uint32_t *__restrict__ medium2,
                                                     uint32_t *__restrict__ large2, uint32_t *__restrict__ tile0_2,
                                                     int tile)
{
    constexpr int CAP = BLOCK * ITEMS;
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int PER = HYB_NBIN / BLOCK; // bins per thread in the scan
    static_assert(HYB_NBIN == 16 * BLOCK, "the scan: 16 bins per thread");
    static_assert(CAP < 65536, "in-bucket slots fit 16 bits");
    __shared__ uint32_t s_cnt[HYB_NBIN + HYB_NBIN / 16];
    __shared__ uint16_t s_low[CAP];
    __shared__ uint32_t s_wsum[WAVES], s_wmax[WAVES], s_top;
    const uint32_t k = blockIdx.x / RADIX, d = blockIdx.x % RADIX;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    if (plan->hybrid != HYB_MSD) {
        // the LSD passes: a third varying byte takes the first overflow
        // pass over one segment of all keys
        if (plan->hybrid == HYB_LSD && plan->hyb_ov_shift[0] < HYB_SKIP && blockIdx.x == 0 && tid == 0) {
            const int64_t n = begin1[RADIX];
            pairs[0] = 0;
            pairs[1] = n;
            tile0_2[0] = 0u;
            large2[0] = 0u;
            sp2->num_large = 1u;
            sp2->num_tiles = uint32_t((n + tile - 1) / tile);
        }
        return;
    }
    if (k >= sp->num_large) return;
    const uint32_t s = large_list[k];
    const int64_t base = begin1[s];
    const uint32_t *__restrict__ first = dstart + size_t(k) * RADIX;
    const int64_t lo = base + first[d];
    const int64_t hi = d + 1 < uint32_t(RADIX) ? base + first[d + 1] : begin1[s + 1];
    const int64_t len64 = hi - lo;
    if (len64 <= 1) return;
    bool over = len64 > CAP;
    const int len = over ? 0 : int(len64);
    // the bins: the 12 bits below level 2's digit (bits hyb_shift2 - 12 ..,
    // at least bit 0), which vary in dense keys as well as in spread ones
    const int bs = max(0, int(plan->hyb_shift2) - 12);
    uint32_t *__restrict__ seg = keys + lo;
    uint32_t v[ITEMS];
    if (!over) {
        // keys striped over the workgroup (coalesced), issued before the LDS set-up
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int o = j * BLOCK + tid;
            v[j] = o < len ? seg[o] : 0u;
        }
        for (int i = tid; i < HYB_NBIN + HYB_NBIN / 16; i += BLOCK) s_cnt[i] = 0u;
        if (tid == 0) s_top = v[0] & 0xFFFF0000u; // key 0: every key of the sub-bucket shares bits 16..31
        __syncthreads();
#pragma unroll
        for (int j = 0; j < ITEMS; ++j)
            if (j * BLOCK + tid < len) atomicAdd(&s_cnt[hyb_slot(int((v[j] >> bs) & (HYB_NBIN - 1)))], 1u);
        __syncthreads();
        // exclusive scan of the bins (thread t owns bins PER t ..) and the largest bin
        uint32_t c[PER];
        uint32_t sum = 0, mx = 0;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            c[i] = s_cnt[hyb_slot(PER * tid + i)];
            sum += c[i];
            mx = max(mx, c[i]);
        }
        uint32_t x = sum;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) mx = max(mx, __shfl_xor(mx, off));
        if (lane == WAVE - 1) s_wsum[wave] = x;
        if (lane == 0) s_wmax[wave] = mx;
        __syncthreads();
        uint32_t run = x - sum, gmax = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            if (w < wave) run += s_wsum[w];
            gmax = max(gmax, s_wmax[w]);
        }
        over = gmax > uint32_t(HYB_MAX_BIN); // the same for every thread
        if (!over) {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                s_cnt[hyb_slot(PER * tid + i)] = run;
                run += c[i];
            }
        }
    }
    if (over) {
        if (tid == 0) {
            const uint32_t i = atomicAdd(&sp2->bad, 1u); // pairs listed so far (no segment is bad here)
            pairs[2 * size_t(i)] = lo;
            pairs[2 * size_t(i) + 1] = hi;
            if (len64 <= SEG_MEDIUM) {
                medium2[atomicAdd(&sp2->num_medium, 1u)] = 2 * i;
            } else {
                tile0_2[2 * size_t(i)] = atomicAdd(&sp2->num_tiles, uint32_t((len64 + tile - 1) / tile));
                large2[atomicAdd(&sp2->num_large, 1u)] = 2 * i;
            }
        }
        return;
    }
    __syncthreads();
    // place each key's low 16 bits at its bin's next free slot (the order
    // inside a bin is the atomics' order; the rank below fixes it)
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        if (j * BLOCK + tid < len) {
            const uint32_t pos = atomicAdd(&s_cnt[hyb_slot(int((v[j] >> bs) & (HYB_NBIN - 1)))], 1u);
            s_low[pos] = uint16_t(v[j]);
        }
    __syncthreads();
    // s_cnt now holds each bin's end; the key at slot p of bin [lb, hb) goes
    // to lb + (keys of the bin below it) + (equal keys at earlier slots):
    // the keys (low 16 bits, slot) before it in one 32-bit order.  When the
    // bins cover every bit below s2 (s2 <= 12), a bin holds equal keys only
    // and slot p is the key's place.
    const uint32_t top = s_top;
    if (plan->hyb_shift2 <= 12u) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int p = j * BLOCK + tid;
            if (p < len) seg[p] = top | s_low[p];
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int p = j * BLOCK + tid;
        if (p < len) {
            const uint32_t xk = s_low[p];
            const int bin = int((xk >> bs) & (HYB_NBIN - 1));
            const uint32_t hb = s_cnt[hyb_slot(bin)];
            const uint32_t lb = bin ? s_cnt[hyb_slot(bin - 1)] : 0u;
            const uint32_t me = (xk << 16) | uint32_t(p);
            uint32_t r = 0;
            for (uint32_t q = lb; q < hb; ++q) r += (((uint32_t(s_low[q]) << 16) | q) < me) ? 1u : 0u;
            seg[lb + r] = top | xk;
        }
    }
}

// ---- Batched segments, MSD-first (hpcs_[dev_]sort_segments_u32 with an
// MSD threshold): segments longer than msd_min keys take pass 0 of the
// segmented passes by their top byte, A -> B (their tile records carry
// SEG_MSD_TILE; passes 1..3 skip them), which leaves 256 buckets per segment
// whose keys share bits 24..31; then each bucket is sorted in LDS from B into
// the same place of A.  Keys alone carry no identity, so the output equals
// the LSD passes'.  Bytes moved per key of such a segment: count 4, top-byte
// pass 8, bucket sort 8 = 20 (the LSD passes: 48).

// The overflow plan cleared and the overflow passes' digit positions
// (bits 0, 8, 16: the buckets share the top byte).  One wave.
static __global__ __launch_bounds__(WAVE) void k_segmsd_init(SegPlan *__restrict__ sp2, uint32_t *__restrict__ shifts)
{
    static_assert(sizeof(SegPlan) / 4 <= WAVE, "one lane per word of the overflow plan");
    const int i = threadIdx.x;
    if (i < int(sizeof(SegPlan) / 4)) reinterpret_cast<uint32_t *>(sp2)[i] = 0u;
    if (i < 3) shifts[i] = 8u * uint32_t(i);
}

// The buckets of the MSD-first segments as (first, end) key positions:
// grid-stride over the large list, one thread per digit; a segment longer
// than msd_min keys takes the next 256 entries of `bounds` (counted in
// sp2->num_small, which the overflow plan does not use).  Bucket d of
// segment s is [b + row[d], b + row[d + 1]) with b = begin[s] and row =
// k_seg_scan's pass-0 offsets of the segment's first tile (the later passes
// leave the row alone); the last one ends at the segment's end.
static __global__ __launch_bounds__(RADIX) void k_segmsd_bounds(const int64_t *__restrict__ begin,
                                                                const SegPlan *__restrict__ sp,
                                                                const uint32_t *__restrict__ seg_tile0,
                                                                const uint32_t *__restrict__ large_list,
                                                                const uint32_t *__restrict__ offsets, int64_t msd_min,
                                                                SegPlan *__restrict__ sp2, uint2 *__restrict__ bounds)
{
    __shared__ uint32_t s_slot;
    const uint32_t d = threadIdx.x, nlarge = sp->num_large;
    for (uint32_t k = blockIdx.x; k < nlarge; k += gridDim.x) {
        const uint32_t s = large_list[k];
        const int64_t b = begin[s], e = begin[s + 1];
        if (e - b <= msd_min) continue; // the LSD passes sorted it
        if (d == 0) s_slot = atomicAdd(&sp2->num_small, 1u);
        __syncthreads();
        const uint32_t *__restrict__ row = offsets + size_t(seg_tile0[s]) * RADIX;
        const int64_t hi = d + 1 < uint32_t(RADIX) ? b + row[d + 1] : e;
        bounds[size_t(s_slot) * RADIX + d] = make_uint2(uint32_t(b + row[d]), uint32_t(hi));
        __syncthreads();
    }
}

// One bucket of len (2 .. CAP) keys in v (striped over the workgroup; every
// key shares bits 24..31) sorted in LDS into dst: counts of the 2^NB bins of
// bits 24 - NB .. 23, their scan, each key placed at its bin's next free
// slot, then each key ranked within its bin by (the bits below the bin,
// slot).  Returns false, having written nothing, when a bin holds more than
// HYB_MAX_BIN keys.  Every thread calls it; it ends after a barrier.
template <int BLOCK, int ITEMS, int NB>
__device__ inline bool segmsd_sort_bucket(const uint32_t (&v)[ITEMS], uint32_t len, uint32_t *__restrict__ dst,
                                          uint32_t *s_cnt, uint32_t *s_key, uint32_t *s_wsum, uint32_t *s_wmax)
{
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int NBIN = 1 << NB;
    constexpr int PER = NBIN / BLOCK; // bins per thread in the scan
    constexpr int BS = 24 - NB;       // the bins' lowest bit
    static_assert(PER >= 1 && PER <= 16, "1 to 16 bins per thread");
    static_assert(BLOCK * ITEMS <= (1 << (8 + NB)), "a slot fits below the compared bits");
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    for (int i = tid; i < NBIN + NBIN / 16; i += BLOCK) s_cnt[i] = 0u;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        if (uint32_t(j * BLOCK + tid) < len) atomicAdd(&s_cnt[hyb_slot(int((v[j] >> BS) & (NBIN - 1)))], 1u);
    __syncthreads();
    uint32_t c[PER];
    uint32_t sum = 0, mx = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        c[i] = s_cnt[hyb_slot(PER * tid + i)];
        sum += c[i];
        mx = max(mx, c[i]);
    }
    uint32_t x = sum;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) mx = max(mx, __shfl_xor(mx, off));
    if (lane == WAVE - 1) s_wsum[wave] = x;
    if (lane == 0) s_wmax[wave] = mx;
    __syncthreads();
    uint32_t run = x - sum, gmax = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        if (w < wave) run += s_wsum[w];
        gmax = max(gmax, s_wmax[w]);
    }
    if (gmax > uint32_t(HYB_MAX_BIN)) return false; // the same for every thread; s_cnt was read before the barrier
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        s_cnt[hyb_slot(PER * tid + i)] = run;
        run += c[i];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        if (uint32_t(j * BLOCK + tid) < len) {
            const uint32_t pos = atomicAdd(&s_cnt[hyb_slot(int((v[j] >> BS) & (NBIN - 1)))], 1u);
            s_key[pos] = v[j];
        }
    __syncthreads();
    // s_cnt holds each bin's end; the key at slot p of bin [lb, hb) goes to
    // lb + the keys of the bin before it in (bits below the bin, slot) order
    // (the bin's keys agree on every higher bit)
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t p = uint32_t(j * BLOCK + tid);
        if (p < len) {
            const uint32_t xk = s_key[p];
            const int bin = int((xk >> BS) & (NBIN - 1));
            const uint32_t hb = s_cnt[hyb_slot(bin)];
            const uint32_t lb = bin ? s_cnt[hyb_slot(bin - 1)] : 0u;
            const uint32_t me = (xk << (8 + NB)) | p;
            uint32_t r = 0;
            for (uint32_t q = lb; q < hb; ++q) r += ((s_key[q] << (8 + NB)) | q) < me ? 1u : 0u;
            dst[lb + r] = xk;
        }
    }
    __syncthreads(); // the next bucket reuses s_cnt and s_key
    return true;
}

// Grid-stride over the buckets g < 256 * sp2->num_small (bounds[g]; lane l
// of every wave fetches the bounds of the workgroup's next 64 buckets at
// once).  A bucket of keysB of at most CAP keys is sorted in LDS into the
// same place of keysA (segmsd_sort_bucket) with 2^10 bins up to 2048 keys
// and 2^11 beyond, so the bins' set-up and scan stay in proportion to the
// keys (and the LDS, 29 KB, leaves room for five workgroups per CU).  Buckets beyond CAP keys or with a bin beyond
// HYB_MAX_BIN keys go to the overflow plan sp2 as k_hyb_local lists them
// (pairs, medium2, large2 with tiles in tile0_2).
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK, BLOCK == 512 ? 8 : 5) void k_segmsd_local(const uint32_t *__restrict__ keysB,
                                                        uint32_t *__restrict__ keysA,
                                                        const uint2 *__restrict__ bounds, SegPlan *__restrict__ sp2,
                                                        int64_t *__restrict__ pairs, uint32_t *__restrict__ medium2,
                                                        uint32_t *__restrict__ large2, uint32_t *__restrict__ tile0_2,
                                                        int tile)
{
    constexpr int CAP = BLOCK * ITEMS;
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int NBIN = 2048;
    static_assert(NBIN <= 16 * BLOCK, "at most 16 bins per thread");
    static_assert(CAP < 65536, "in-bucket slots fit 16 bits");
    __shared__ uint32_t s_cnt[NBIN + NBIN / 16];
    __shared__ uint32_t s_key[CAP];
    __shared__ uint32_t s_wsum[WAVES], s_wmax[WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const uint32_t total = sp2->num_small * uint32_t(RADIX), G = gridDim.x;
    // an overflow bucket: listed by thread 0
    auto overflow = [&](uint32_t lo, uint32_t hi) {
        if (tid != 0) return;
        const uint32_t len = hi - lo;
        const uint32_t i = atomicAdd(&sp2->bad, 1u); // pairs listed so far (no segment is bad here)
        pairs[2 * size_t(i)] = int64_t(lo);
        pairs[2 * size_t(i) + 1] = int64_t(hi);
        if (len <= uint32_t(SEG_MEDIUM)) {
            medium2[atomicAdd(&sp2->num_medium, 1u)] = 2 * i;
        } else {
            tile0_2[2 * size_t(i)] = atomicAdd(&sp2->num_tiles, (len + uint32_t(tile) - 1) / uint32_t(tile));
            large2[atomicAdd(&sp2->num_large, 1u)] = 2 * i;
        }
    };
    uint2 nb = make_uint2(0u, 0u); // lane l: the bounds of iteration (it & ~63) + l
    uint32_t it = 0;
    for (uint32_t g = blockIdx.x; g < total; g += G, ++it) {
        if ((it & (WAVE - 1)) == 0) {
            const uint32_t gl = g + uint32_t(lane) * G;
            nb = gl < total ? bounds[gl] : make_uint2(0u, 0u);
        }
        const uint32_t lo = __builtin_amdgcn_readlane(nb.x, int(it & (WAVE - 1)));
        const uint32_t hi = __builtin_amdgcn_readlane(nb.y, int(it & (WAVE - 1)));
        const uint32_t len = hi - lo;
        if (len <= 1u) {
            if (len == 1u && tid == 0) keysA[lo] = keysB[lo];
            continue;
        }
        if (len > uint32_t(CAP)) {
            overflow(lo, hi);
            continue;
        }
        const uint32_t *__restrict__ src = keysB + lo;
        uint32_t v[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t o = uint32_t(j * BLOCK + tid);
            v[j] = o < len ? src[o] : 0u;
        }
        const bool done = len <= 2048u
                              ? segmsd_sort_bucket<BLOCK, ITEMS, 10>(v, len, keysA + lo, s_cnt, s_key, s_wsum, s_wmax)
                              : segmsd_sort_bucket<BLOCK, ITEMS, 11>(v, len, keysA + lo, s_cnt, s_key, s_wsum, s_wmax);
        if (!done) overflow(lo, hi);
    }
}

} // namespace hpcs
