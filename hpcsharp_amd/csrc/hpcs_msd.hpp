// hpcs_msd.hpp -- one MSD split, then one workgroup per bucket (keys only).
//
// For arrays of up to a few million keys the LSD passes are launch-bound:
// each pass is a count, a scan and a scatter over few tiles.  Here instead:
//   k_or_and      OR(keys) and OR(~keys): the bits in which the keys differ
//   k_msd_hist    the 256-bin histogram of the 8-bit digit at the top of
//                 those bits (shift = highest varying bit - 7, at least 0),
//                 so that keys spread over the buckets even when they use
//                 only part of the key range
//   k_plan_msd    exclusive starts of the buckets
//   k_onesweep    (pass -1) one stable scatter of A into B by that digit,
//                 keys left in radix-key form
//   k_sort_buckets  workgroup b sorts bucket b: up to its tile in LDS (LSD
//                 passes over the digit positions that vary inside the
//                 bucket), larger buckets by streaming stable passes through
//                 global memory; the result goes, decoded, to its place in A
// Stability: the split is stable and every bucket sort is stable, so equal
// keys keep their input order (the LSD contract of RadixSortLsd.cs:127-202).
#pragma once

#include "hpcs_kernels.hpp"
#include "hpcs_segments.hpp"

namespace hpcs {

constexpr int MSD_BLOCK = 256; // k_or_and / k_msd_hist

template <typename K> __device__ inline int highest_bit(K x);
template <> __device__ inline int highest_bit<uint32_t>(uint32_t x) { return x ? 31 - __clz(x) : -1; }
template <> __device__ inline int highest_bit<uint64_t>(uint64_t x) { return x ? 63 - __clzll(x) : -1; }

// OR and OR-of-complement of the radix keys into one of ORSLOTS slot pairs.
template <typename K, int FLOATK>
__global__ __launch_bounds__(MSD_BLOCK) void k_or_and(const K *__restrict__ keys, uint32_t n, K *__restrict__ orslots)
{
    constexpr int WAVES = MSD_BLOCK / WAVE;
    __shared__ K s_or[WAVES], s_nor[WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    K vor = 0, vnor = 0;
    for (uint32_t i = blockIdx.x * MSD_BLOCK + tid; i < n; i += gridDim.x * MSD_BLOCK) {
        const K k = FLOATK ? key_enc<K, FLOATK>(keys[i]) : keys[i];
        vor |= k;
        vnor |= ~k;
    }
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        vor |= __shfl_xor(vor, off);
        vnor |= __shfl_xor(vnor, off);
    }
    if (lane == 0) {
        s_or[wave] = vor;
        s_nor[wave] = vnor;
    }
    __syncthreads();
    if (tid == 0) {
        K a = 0, b = 0;
        for (int w = 0; w < WAVES; ++w) {
            a |= s_or[w];
            b |= s_nor[w];
        }
        const int slot = blockIdx.x % ORSLOTS;
        atomicOr(&orslots[slot * OR_STRIDE<K>], a);
        atomicOr(&orslots[slot * OR_STRIDE<K> + 1], b);
    }
}

template <typename K> __device__ inline K varying_bits(const K *__restrict__ orslots)
{
    K a = 0, b = 0;
    for (int s = 0; s < ORSLOTS; ++s) {
        a |= orslots[s * OR_STRIDE<K>];
        b |= orslots[s * OR_STRIDE<K> + 1];
    }
    return a & b;
}

template <typename K> __device__ inline int msd_shift_of(K varying)
{
    const int hb = highest_bit<K>(varying);
    return hb >= RADIX_BITS - 1 ? hb - (RADIX_BITS - 1) : 0;
}

// Histogram of the split digit into plan->msd_start (zero-filled before).
template <typename K, int FLOATK>
__global__ __launch_bounds__(MSD_BLOCK) void k_msd_hist(const K *__restrict__ keys, uint32_t n,
                                                        const K *__restrict__ orslots, DevPlan *__restrict__ plan)
{
    __shared__ uint32_t sh[RADIX];
    __shared__ int s_shift;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    for (int i = tid; i < RADIX; i += MSD_BLOCK) sh[i] = 0u;
    if (tid == 0) s_shift = msd_shift_of<K>(varying_bits<K>(orslots));
    __syncthreads();
    const int shift = s_shift;
    for (uint32_t i = blockIdx.x * MSD_BLOCK + tid; i < n; i += gridDim.x * MSD_BLOCK) {
        const K k = FLOATK ? key_enc<K, FLOATK>(keys[i]) : keys[i];
        const uint32_t d = KeyOps<K>::digit(k, shift);
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        const uint32_t active = uint32_t(__popcll(__ballot(true))); // lanes still in the loop
        if (__all(d == d0)) {
            if (lane == 0) atomicAdd(&sh[d0], active);
        } else {
            atomicAdd(&sh[d], 1u);
        }
    }
    __syncthreads();
    for (int i = tid; i < RADIX; i += MSD_BLOCK)
        if (sh[i]) atomicAdd(&plan->msd_start[i], sh[i]);
}

// Bucket starts (exclusive scan of the split digit's counts), the split's
// shift and its scatter flags (from A, keys transformed on load and left in
// radix-key form in B).  Every key equal: the split is trivial and so is
// every bucket sort.
template <typename K>
__global__ __launch_bounds__(RADIX) void k_plan_msd(const K *__restrict__ orslots, DevPlan *__restrict__ plan,
                                                    uint32_t n, int floatk)
{
    __shared__ uint32_t s[RADIX];
    const int b = threadIdx.x;
    const K varying = varying_bits<K>(orslots);
    const uint32_t c = plan->msd_start[b];
    s[b] = c;
    __syncthreads();
    for (int off = 1; off < RADIX; off <<= 1) {
        const uint32_t add = b >= off ? s[b - off] : 0u;
        __syncthreads();
        s[b] += add;
        __syncthreads();
    }
    plan->msd_start[b] = s[b] - c;
    if (b == 0) {
        plan->msd_shift = uint32_t(msd_shift_of<K>(varying));
        plan->msd_flags = (varying == 0 || n <= 1) ? PF_TRIVIAL : (floatk ? PF_XIN : 0u);
    }
}

// Workgroup b sorts bucket b of B (keys in radix-key form) into A (decoded).
// Buckets of up to BLOCK * ITEMS keys: one read, the LSD passes in LDS over
// the digit positions that vary inside the bucket, one write.  Larger
// buckets: per digit position that varies inside the bucket, a count of the
// bucket and a stable scatter of it tile by tile (in-tile ranking as in
// LDS), ping-ponging between the bucket's ranges of B and A, then a final
// decoded copy into A.
template <typename K, int FLOATK, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_sort_buckets(K *__restrict__ A, K *__restrict__ B, uint32_t n,
                                                        const DevPlan *__restrict__ plan)
{
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int P = int(sizeof(K));
    __shared__ uint32_t s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[WAVES][RADIX]; // LDS-OR match ranking
    __shared__ uint32_t s_scan[RADIX / WAVE];
    __shared__ uint32_t s_start[RADIX];
    __shared__ uint32_t s_off[RADIX];
    __shared__ K s_keys[TILE];
    __shared__ K s_or[WAVES], s_nor[WAVES];
    if (plan->msd_flags & PF_TRIVIAL) return; // every key equal: A is sorted
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    const uint32_t bkt = blockIdx.x;
    const uint32_t beg = plan->msd_start[bkt];
    const uint32_t end = bkt + 1 < RADIX ? plan->msd_start[bkt + 1] : n;
    if (end <= beg) return;
    const uint32_t len = end - beg;

    // bits that vary inside the bucket
    K vor = 0, vnor = 0;
    K v[ITEMS];
    if (len <= uint32_t(TILE)) {
        const int wave_valid = int(len) > wave * WAVE_TILE ? min(WAVE_TILE, int(len) - wave * WAVE_TILE) : 0;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int o = j * WAVE + lane;
            v[j] = ~K(0);
            if (o < wave_valid) {
                v[j] = B[beg + uint32_t(wave * WAVE_TILE + o)];
                vor |= v[j];
                vnor |= ~v[j];
            }
        }
    } else {
        for (uint32_t i = tid; i < len; i += BLOCK) {
            const K k = B[beg + i];
            vor |= k;
            vnor |= ~k;
        }
    }
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        vor |= __shfl_xor(vor, off);
        vnor |= __shfl_xor(vnor, off);
    }
    if (lane == 0) {
        s_or[wave] = vor;
        s_nor[wave] = vnor;
    }
    __syncthreads();
    K a = 0, c = 0;
    for (int w = 0; w < WAVES; ++w) {
        a |= s_or[w];
        c |= s_nor[w];
    }
    const K varying = a & c;

    if (len <= uint32_t(TILE)) {
        const int wave_valid = int(len) > wave * WAVE_TILE ? min(WAVE_TILE, int(len) - wave * WAVE_TILE) : 0;
        bool staged = false;
        for (int pass = 0; pass < P; ++pass) {
            if (((varying >> (8 * pass)) & K(0xFF)) == 0) continue;
            if (staged) {
#pragma unroll
                for (int j = 0; j < ITEMS; ++j) {
                    const int o = j * WAVE + lane;
                    if (o < wave_valid) v[j] = s_keys[wave * WAVE_TILE + o];
                }
                __syncthreads();
            }
            tile_rank_stage_ldsor<K, BLOCK, ITEMS>(v, wave_valid, 8 * pass, s_cnt, s_match, s_scan, s_start, s_keys);
            staged = true;
        }
        if (staged) {
            for (uint32_t j = tid; j < len; j += BLOCK)
                A[beg + j] = FLOATK ? key_dec<K, FLOATK>(s_keys[j]) : s_keys[j];
        } else { // one key value in the bucket: copy it over in order
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const int o = j * WAVE + lane;
                if (o < wave_valid) A[beg + uint32_t(wave * WAVE_TILE + o)] = FLOATK ? key_dec<K, FLOATK>(v[j]) : v[j];
            }
        }
        return;
    }

    // ---- a bucket larger than the tile: streaming stable passes ----
    bool in_a = false; // the bucket's keys sit in B
    for (int pass = 0; pass < P; ++pass) {
        if (((varying >> (8 * pass)) & K(0xFF)) == 0) continue;
        const int shift = 8 * pass;
        const K *__restrict__ src = (in_a ? A : B) + beg;
        K *__restrict__ dst = (in_a ? B : A) + beg;
        // count the bucket's digits, then their exclusive starts
        for (int i = tid; i < RADIX; i += BLOCK) s_off[i] = 0u;
        __syncthreads();
        for (uint32_t i = tid; i < len; i += BLOCK) atomicAdd(&s_off[KeyOps<K>::digit(src[i], shift)], 1u);
        __syncthreads();
        uint32_t cnt = 0, xs = 0;
        if (tid < RADIX) {
            cnt = s_off[tid];
            xs = cnt;
#pragma unroll
            for (int off = 1; off < WAVE; off <<= 1) {
                const uint32_t y = __shfl_up(xs, off);
                if (lane >= off) xs += y;
            }
            if (lane == WAVE - 1) s_scan[wave] = xs;
        }
        __syncthreads();
        if (tid < RADIX) {
            uint32_t base = xs - cnt;
            for (int w = 0; w < wave; ++w) base += s_scan[w];
            s_off[tid] = base;
        }
        __syncthreads();
        // scatter tile by tile in input order: in-tile stable ranking, each
        // digit's run of the tile at that digit's running offset
        for (uint32_t t0 = 0; t0 < len; t0 += uint32_t(TILE)) {
            const uint32_t tl = min(uint32_t(TILE), len - t0);
            const int wave_valid = int(tl) > wave * WAVE_TILE ? min(WAVE_TILE, int(tl) - wave * WAVE_TILE) : 0;
            K u[ITEMS];
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const int o = j * WAVE + lane;
                u[j] = o < wave_valid ? src[t0 + uint32_t(wave * WAVE_TILE + o)] : ~K(0);
            }
            tile_rank_stage_ldsor<K, BLOCK, ITEMS>(u, wave_valid, shift, s_cnt, s_match, s_scan, s_start, s_keys);
            for (uint32_t j = tid; j < tl; j += BLOCK) {
                const K k = s_keys[j];
                const uint32_t d = KeyOps<K>::digit(k, shift);
                dst[s_off[d] + (j - s_start[d])] = k;
            }
            __syncthreads();
            if (tid < RADIX) {
                const uint32_t next = tid + 1 < RADIX ? s_start[tid + 1] : tl;
                s_off[tid] += next - s_start[tid];
            }
            __syncthreads();
        }
        in_a = !in_a;
    }
    // the sorted bucket into A, decoded
    if (!in_a || FLOATK) {
        const K *__restrict__ src = (in_a ? A : B) + beg;
        for (uint32_t i = tid; i < len; i += BLOCK) {
            const K k = src[i];
            A[beg + i] = FLOATK ? key_dec<K, FLOATK>(k) : k;
        }
    }
}

} // namespace hpcs
