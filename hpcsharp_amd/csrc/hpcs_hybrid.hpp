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

} // namespace hpcs
