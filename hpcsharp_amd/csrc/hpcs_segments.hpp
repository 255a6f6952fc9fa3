// hpcs_segments.hpp -- batched sort of many uint32 segments in one fixed
// sequence of launches (the GPU form of calling the Experimental
// SortRadixLsd(uint[], startIndex, length), RadixSortLsd.cs:1409-1457, once
// per segment).
//
// Segment s is keys[begin[s], begin[s+1]).  Everything is planned on the
// device, so the launch sequence does not depend on the segment lengths and
// can be captured once in a HIP graph and replayed:
//   k_seg_plan     classify: segments of at most SEG_SMALL keys go to the
//                  small list, those of at most small_max (SEG_SMALL or
//                  SEG_MEDIUM, a launch argument) to the medium list, both
//                  sorted by k_seg_small; larger ones get a range of tiles of
//                  `tile` keys (a launch argument: the tile shape's BLOCK x
//                  ITEMS; a tile never crosses a segment)
//   k_seg_tiles    one 16-byte record per tile (its keys, its segment)
//   k_seg_small    one workgroup per small (256 threads) or medium (1024
//                  threads) segment: the LSD passes entirely in LDS (one HBM
//                  read and one write per key)
//   per pass p of the large segments:
//     k_seg_counts   per tile, the digit counts (16-bit: a tile holds fewer
//                    than 65536 keys)
//     k_seg_scan     per segment, digit starts (segment base + exclusive scan
//                    of the segment's digit totals) and per-tile offsets
//                    (32-bit, positions within the segment)
//     k_seg_scatter  rank (RANK as the plain sort's scatter), stage in LDS,
//                    write each digit run contiguously
// Four passes leave every large segment back in the caller's buffer.
#pragma once
#include "hpcs_kernels.hpp"

namespace hpcs {

// The small list's workgroup (and the default tile of the large segments)
constexpr int SEG_BLOCK = 256;
constexpr int SEG_ITEMS = 16;
constexpr int SEG_TILE = SEG_BLOCK * SEG_ITEMS; // 4096 keys
constexpr int SEG_SMALL = SEG_TILE;             // largest segment of the small list (256 threads)
constexpr int SEG_MIN_TILE = 2048;              // smallest tile of a shape that serves the large segments
constexpr int SEG_MEDIUM_BLOCK = 1024;
constexpr int SEG_MEDIUM = SEG_MEDIUM_BLOCK * SEG_ITEMS; // largest segment one workgroup sorts in LDS
// Tile record flag (in .w, beside the large-list index): the tile's segment
// takes the MSD-first form, whose only segmented pass is pass 0 by the top
// byte (bits 24..31); passes 1..3 skip its tiles.
constexpr uint32_t SEG_MSD_TILE = 0x80000000u;

struct SegPlan {
    uint32_t num_tiles;  // tiles of the large segments
    uint32_t num_small;  // small segments
    uint32_t num_large;  // large segments
    uint32_t bad;        // segments skipped: end < begin, outside [0, n), or starting before
                         // the end of an earlier segment (overlap)
    uint32_t num_medium; // medium segments (SEG_SMALL < keys <= small_max)
    uint32_t pad[59];
};

// Ballot ranking of one tile held lane-striped in registers (wave w, item i,
// lane l = key w*64*ITEMS + i*64 + l): on return every valid key (and, with
// HAS_VALS, its value) has been stored to s_keys (s_vals) at its stable
// in-tile position sorted by the digit at `shift`, and s_start[d] holds the
// tile's first position of digit d.
template <typename K, int BLOCK, int ITEMS, bool HAS_VALS = false>
__device__ inline void tile_rank_stage(const K (&keys)[ITEMS], int wave_valid, int shift,
                                       uint32_t (*s_cnt)[RADIX], uint32_t *s_scan, uint32_t *s_start, K *s_keys,
                                       const uint32_t *vals = nullptr, uint32_t *s_vals = nullptr)
{
    constexpr int WAVES = BLOCK / WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    for (int q = tid; q < WAVES * RADIX; q += BLOCK) (&s_cnt[0][0])[q] = 0u;
    __syncthreads();
    uint32_t *cnt = s_cnt[wave];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = KeyOps<K>::digit(keys[j], shift);
        const bool valid = j * WAVE + lane < wave_valid;
        const uint64_t peers = match_digit8(d) & __ballot(valid);
        const uint32_t below = count_below(peers);
        const uint32_t old = cnt[d];
        if (valid && below == 0) cnt[d] = old + uint32_t(__popcll(peers));
        rank[j] = old + below;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    uint32_t tile_cnt = 0, tile_start = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = tile_cnt;
            tile_cnt += c;
        }
        uint32_t xs = tile_cnt;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(xs, off);
            if (lane >= off) xs += y;
        }
        if (lane == WAVE - 1) s_scan[wave] = xs;
        tile_start = xs - tile_cnt;
    }
    __syncthreads();
    if (tid < RADIX) {
        for (int w = 0; w < wave; ++w) tile_start += s_scan[w];
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s_cnt[w][tid] += tile_start;
        s_start[tid] = tile_start;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        if (j * WAVE + lane < wave_valid) {
            const uint32_t at = cnt[KeyOps<K>::digit(keys[j], shift)] + rank[j];
            s_keys[at] = keys[j];
            if constexpr (HAS_VALS) s_vals[at] = vals[j];
        }
    }
    __syncthreads();
}

// As tile_rank_stage, ranking with LDS-OR match masks: per-wave
// counts by LDS atomics first, then each lane ORs its bit into its digit's
// mask, reads the mask back and stages its key at base + lanes below it.
template <typename K, int BLOCK, int ITEMS, bool HAS_VALS = false>
__device__ inline void tile_rank_stage_ldsor(const K (&keys)[ITEMS], int wave_valid, int shift,
                                             uint32_t (*s_cnt)[RADIX], uint64_t (*s_match)[RADIX],
                                             uint32_t *s_scan, uint32_t *s_start, K *s_keys,
                                             const uint32_t *vals = nullptr, uint32_t *s_vals = nullptr)
{
    constexpr int WAVES = BLOCK / WAVE;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    for (int q = tid; q < WAVES * RADIX; q += BLOCK) {
        (&s_cnt[0][0])[q] = 0u;
        (&s_match[0][0])[q] = 0ull;
    }
    __syncthreads();
    uint32_t *cnt = s_cnt[wave];
    const bool skewed = wave_skewed(KeyOps<K>::digit(keys[0], shift), lane < wave_valid,
                                    KeyOps<K>::digit(keys[ITEMS / 2], shift), (ITEMS / 2) * WAVE + lane < wave_valid);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
        count_item(cnt, KeyOps<K>::digit(keys[j], shift), j * WAVE + lane < wave_valid, lane, skewed);
    __syncthreads();
    uint32_t tile_cnt = 0, tile_start = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = tile_cnt;
            tile_cnt += c;
        }
        uint32_t xs = tile_cnt;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(xs, off);
            if (lane >= off) xs += y;
        }
        if (lane == WAVE - 1) s_scan[wave] = xs;
        tile_start = xs - tile_cnt;
    }
    __syncthreads();
    if (tid < RADIX) {
        for (int w = 0; w < wave; ++w) tile_start += s_scan[w];
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s_cnt[w][tid] += tile_start;
        s_start[tid] = tile_start;
    }
    __syncthreads();
    uint32_t *match32 = reinterpret_cast<uint32_t *>(s_match[wave]);
    auto rank_loop = [&](auto peel_c) {
        constexpr bool PEEL = decltype(peel_c)::value;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool valid = j * WAVE + lane < wave_valid;
            const uint32_t at =
                ldsor_rank_item<PEEL>(KeyOps<K>::digit(keys[j], shift), valid, lane, cnt, match32, s_match[wave]);
            if (valid) {
                s_keys[at] = keys[j];
                if constexpr (HAS_VALS) s_vals[at] = vals[j];
            }
            __builtin_amdgcn_wave_barrier();
        }
    };
    if (skewed) rank_loop(std::true_type{});
    else rank_loop(std::false_type{});
    __syncthreads();
}

// One block of 1024 threads.  begin has nseg + 1 entries.  Per segment: its
// first tile (seg_tile0) and its place in the small or large list.
// A segment is sorted only if 0 <= begin <= end <= n and it starts at or
// after the end of every earlier segment with valid bounds (a block-wide
// running maximum of the ends), so the sorted segments never overlap, their
// keys total at most n and their tiles at most n / TILE + nseg, the bound the
// host sized the tile records and counts by (max_tiles; a plan beyond it
// sorts no large segment).
static __global__ __launch_bounds__(1024) void k_seg_plan(const int64_t *__restrict__ begin, int nseg, int64_t n,
                                                   SegPlan *__restrict__ sp, uint32_t *__restrict__ seg_tile0,
                                                   uint32_t *__restrict__ small_list,
                                                   uint32_t *__restrict__ medium_list,
                                                   uint32_t *__restrict__ large_list, uint64_t max_tiles,
                                                   int small_max, int tile, int min_small)
{
    __shared__ uint64_t s_wsum[16];
    __shared__ int64_t s_wmax[16];
    __shared__ uint32_t s_carry[4];
    __shared__ int64_t s_end_carry; // largest valid end of the earlier chunks
    __shared__ uint32_t s_bad;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    if (tid < 4) s_carry[tid] = 0u;
    if (tid == 0) {
        s_bad = 0u;
        s_end_carry = 0;
    }
    __syncthreads();
    for (int base = 0; base < nseg; base += 1024) {
        const int s = base + tid;
        uint32_t tiles = 0, small = 0, medium = 0;
        const int64_t b = s < nseg ? begin[s] : 0, e = s < nseg ? begin[s + 1] : 0;
        const bool bounds_ok = s < nseg && b >= 0 && e >= b && e <= n;
        // exclusive running maximum of the valid ends before segment s
        int64_t m = bounds_ok ? e : 0;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const int64_t y = __shfl_up(m, off);
            if (lane >= off) m = max(m, y);
        }
        if (lane == WAVE - 1) s_wmax[wave] = m;
        __syncthreads();
        int64_t before = s_end_carry;
        for (int w = 0; w < wave; ++w) before = max(before, s_wmax[w]);
        int64_t prev_lane = __shfl_up(m, 1);
        if (lane > 0) before = max(before, prev_lane);
        __syncthreads();
        if (tid == 1023) s_end_carry = max(before, bounds_ok ? e : int64_t(0));
        if (s < nseg) {
            if (!bounds_ok || b < before) {
                atomicAdd(&s_bad, 1u);
            } else {
                const int64_t len = e - b;
                if (len > small_max) tiles = uint32_t((len + tile - 1) / tile);
                else if (len > SEG_SMALL) medium = 1;
                else if (len >= min_small) small = 1;
            }
        }
        const uint32_t large = tiles ? 1u : 0u;
        // one block-wide exclusive scan of (tiles, small, large, medium)
        // packed in 64 bits: within a chunk of 1024 segments the tiles sum to
        // at most n / SEG_MIN_TILE + 1024 < 2^22 (the sorted segments do not
        // overlap and n < 2^32) and each list count to at most 1024 < 2^11
        constexpr int SH[4] = {0, 22, 33, 44};
        constexpr uint64_t MASK[4] = {(1ull << 22) - 1, (1ull << 11) - 1, (1ull << 11) - 1, (1ull << 11) - 1};
        const uint64_t v = uint64_t(tiles) | (uint64_t(small) << SH[1]) | (uint64_t(large) << SH[2]) |
                           (uint64_t(medium) << SH[3]);
        uint64_t x = v;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint64_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == WAVE - 1) s_wsum[wave] = x;
        __syncthreads();
        uint64_t add = 0;
        for (int w = 0; w < wave; ++w) add += s_wsum[w];
        const uint64_t ex = add + x - v, inc = add + x;
        uint32_t excl[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) excl[k] = s_carry[k] + uint32_t((ex >> SH[k]) & MASK[k]);
        __syncthreads();
        if (tid == 1023)
#pragma unroll
            for (int k = 0; k < 4; ++k) s_carry[k] += uint32_t((inc >> SH[k]) & MASK[k]);
        __syncthreads();
        if (s < nseg) {
            seg_tile0[s] = excl[0];
            if (small) small_list[excl[1]] = uint32_t(s);
            if (large) large_list[excl[2]] = uint32_t(s);
            if (medium) medium_list[excl[3]] = uint32_t(s);
        }
    }
    __syncthreads();
    if (tid == 0) {
        const bool fits = uint64_t(s_carry[0]) <= max_tiles; // holds for non-overlapping segments
        sp->num_tiles = fits ? s_carry[0] : 0u;
        sp->num_small = s_carry[1];
        sp->num_large = fits ? s_carry[2] : 0u;
        sp->num_medium = s_carry[3];
        sp->bad = s_bad + (fits ? 0u : s_carry[2]);
    }
}

// One record per tile of the large segments: {first key, segment's first key,
// keys in the tile, the segment's place in the large list}, so the pass
// kernels find a tile with one load.  With msd_min > 0, tiles of segments
// longer than msd_min keys carry SEG_MSD_TILE.
// Grid-stride over the large segments, the segment's tiles across the threads.
static __global__ __launch_bounds__(256) void k_seg_tiles(const int64_t *__restrict__ begin,
                                                   const SegPlan *__restrict__ sp,
                                                   const uint32_t *__restrict__ seg_tile0,
                                                   const uint32_t *__restrict__ large_list,
                                                   uint4 *__restrict__ tinfo, int tile, int64_t msd_min)
{
    const uint32_t nlarge = sp->num_large;
    for (uint32_t k = blockIdx.x; k < nlarge; k += gridDim.x) {
        const uint32_t s = large_list[k];
        const int64_t b = begin[s], len = begin[s + 1] - b;
        const uint32_t t0 = seg_tile0[s];
        const uint32_t nt = uint32_t((len + tile - 1) / tile);
        const uint32_t w = k | (msd_min > 0 && len > msd_min ? SEG_MSD_TILE : 0u);
        for (uint32_t t = threadIdx.x; t < nt; t += blockDim.x) {
            const int64_t tb = int64_t(t) * tile;
            tinfo[t0 + t] = make_uint4(uint32_t(b + tb), uint32_t(b), uint32_t(min(int64_t(tile), len - tb)), w);
        }
    }
}

// One workgroup per segment of a list (grid-stride): the LSD passes in LDS.
// BLOCK = SEG_BLOCK: the small list (at most SEG_SMALL keys); BLOCK =
// SEG_MEDIUM_BLOCK: the medium list (at most SEG_MEDIUM keys, one workgroup
// per CU by its LDS).  Digit positions where every key of the segment agrees
// are skipped.  The sorted segment goes to the same place of keys_out, which
// is keys or another buffer (then every listed segment is written, sorted or
// not).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_seg_small(const uint32_t *keys, uint32_t *keys_out,
                                                     const int64_t *__restrict__ begin,
                                                     const SegPlan *__restrict__ sp,
                                                     const uint32_t *__restrict__ list)
{
    static_assert(BLOCK == SEG_BLOCK || BLOCK == SEG_MEDIUM_BLOCK, "the small or the medium list");
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int WAVE_TILE = WAVE * SEG_ITEMS;
    constexpr int TILE = BLOCK * SEG_ITEMS;
    __shared__ uint32_t s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[WAVES][RADIX];
    __shared__ uint32_t s_scan[RADIX / WAVE];
    __shared__ uint32_t s_start[RADIX];
    __shared__ uint32_t s_keys[TILE];
    __shared__ uint32_t s_or[WAVES], s_nor[WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    const uint32_t nlist = BLOCK == SEG_BLOCK ? sp->num_small : sp->num_medium;
    for (uint32_t k = blockIdx.x; k < nlist; k += gridDim.x) {
        const uint32_t s = list[k];
        const int64_t b0 = begin[s];
        const int len = int(begin[s + 1] - b0); // at most TILE (k_seg_plan)
        const uint32_t *seg = keys + b0;
        uint32_t *seg_out = keys_out + b0;
        const int wave_valid = len > wave * WAVE_TILE ? min(WAVE_TILE, len - wave * WAVE_TILE) : 0;
        uint32_t v[SEG_ITEMS];
        uint32_t vor = 0, vnor = 0;
#pragma unroll
        for (int j = 0; j < SEG_ITEMS; ++j) {
            const int o = j * WAVE + lane;
            v[j] = o < wave_valid ? seg[wave * WAVE_TILE + o] : ~0u;
            if (o < wave_valid) {
                vor |= v[j];
                vnor |= ~v[j];
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
        uint32_t varying = 0, a = 0, c = 0;
        for (int w = 0; w < WAVES; ++w) {
            a |= s_or[w];
            c |= s_nor[w];
        }
        varying = a & c;
        bool staged = false; // keys currently in s_keys in sorted order
        for (int pass = 0; pass < 4; ++pass) {
            if (((varying >> (8 * pass)) & 0xFFu) == 0u) continue;
            if (staged) {
#pragma unroll
                for (int j = 0; j < SEG_ITEMS; ++j) {
                    const int o = j * WAVE + lane;
                    v[j] = o < wave_valid ? s_keys[wave * WAVE_TILE + o] : ~0u;
                }
                __syncthreads();
            }
            tile_rank_stage_ldsor<uint32_t, BLOCK, SEG_ITEMS>(v, wave_valid, 8 * pass, s_cnt, s_match, s_scan,
                                                              s_start, s_keys);
            staged = true;
        }
        if (staged) {
            for (int j = tid; j < len; j += BLOCK) seg_out[j] = s_keys[j];
        } else if (keys_out != keys) {
#pragma unroll
            for (int j = 0; j < SEG_ITEMS; ++j) {
                const int o = j * WAVE + lane;
                if (o < wave_valid) seg_out[wave * WAVE_TILE + o] = v[j];
            }
        }
        __syncthreads();
    }
}

// One workgroup per tile (BLOCK x ITEMS keys) of the large segments (the
// grid may exceed the tile count; spare workgroups return): digit counts of
// pass `pass`.  STRIDE: a grid smaller than the tile count walks the tiles
// grid-stride (the overflow stages' launches, which are mostly empty; a
// separate instantiation, since the loop costs the one-tile form registers
// and time).
template <int BLOCK, int ITEMS, bool PEEL>
__device__ __forceinline__ void seg_counts_tile(uint32_t t, const uint32_t *__restrict__ keysA,
                                                const uint32_t *__restrict__ keysB, int pass,
                                                const uint4 *__restrict__ tinfo, uint16_t *__restrict__ counts,
                                                const uint32_t *__restrict__ dshift)
{
    static_assert(BLOCK * ITEMS < 65536, "a tile's digit counts fit 16 bits");
    __shared__ uint32_t sh[RADIX];
    const uint32_t *__restrict__ kin = (pass & 1) ? keysB : keysA;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    int shift = dshift ? int(*dshift) : 8 * pass; // the digit's bit; pass picks the buffers
    if (shift >= 32) return;                      // a pass the plan skips
    const uint4 ti = tinfo[t];
    if (ti.w & SEG_MSD_TILE) { // MSD-first segment: pass 0 by the top byte only
        if (pass != 0) return;
        shift = 24;
    }
    const uint32_t tb = ti.x;
    const int cnt = int(ti.z);
    for (int i = tid; i < RADIX; i += BLOCK) sh[i] = 0u;
    __syncthreads();
    uint32_t k[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int o = j * BLOCK + tid;
        k[j] = o < cnt ? kin[size_t(tb) + o] : 0u;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const uint32_t d = (k[j] >> shift) & 0xFFu;
        if constexpr (PEEL) {
            count_digit(sh, d, j * BLOCK + tid < cnt, lane);
        } else {
            // one digit for the whole wave: one add
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            if (__all(j * BLOCK + tid < cnt && d == d0)) {
                if (lane == 0) atomicAdd(&sh[d0], uint32_t(WAVE));
            } else if (j * BLOCK + tid < cnt) {
                atomicAdd(&sh[d], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < RADIX; i += BLOCK) counts[size_t(t) * RADIX + i] = uint16_t(sh[i]);
}

template <int BLOCK, int ITEMS, bool PEEL = true, bool STRIDE = false>
__global__ __launch_bounds__(BLOCK) void k_seg_counts(const uint32_t *__restrict__ keysA,
                                                      const uint32_t *__restrict__ keysB, int pass,
                                                      const SegPlan *__restrict__ sp, const uint4 *__restrict__ tinfo,
                                                      uint16_t *__restrict__ counts,
                                                      const uint32_t *__restrict__ dshift)
{
    const uint32_t nt = sp->num_tiles;
    if constexpr (!STRIDE) {
        if (blockIdx.x < nt) seg_counts_tile<BLOCK, ITEMS, PEEL>(blockIdx.x, keysA, keysB, pass, tinfo, counts, dshift);
        return;
    }
#pragma nounroll
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        seg_counts_tile<BLOCK, ITEMS, PEEL>(t, keysA, keysB, pass, tinfo, counts, dshift);
        if (t + gridDim.x < nt) __syncthreads(); // the next tile clears the counters
    }
}

// Grid-stride over the large segments, one workgroup each: digit starts of
// the segment and the per-tile offsets (from the 16-bit counts).  The
// SEG_SCAN_GROUPS x SEG_SCAN_PAIRS threads split the segment's tiles into
// SEG_SCAN_GROUPS consecutive parts; thread (g, q) owns digits 2q and 2q + 1
// of part g (one 32-bit load of two counts per tile, one 8-byte store of two
// offsets), so the serial walk over a part is an eighth of the segment.
constexpr int SEG_SCAN_GROUPS = 8;
constexpr int SEG_SCAN_PAIRS = RADIX / 2;
static __global__ __launch_bounds__(SEG_SCAN_GROUPS * SEG_SCAN_PAIRS) void k_seg_scan(const int64_t *__restrict__ begin,
                                                                              const SegPlan *__restrict__ sp,
                                                                              const uint32_t *__restrict__ seg_tile0,
                                                                              const uint32_t *__restrict__ large_list,
                                                                              const uint16_t *__restrict__ counts,
                                                                              uint32_t *__restrict__ offsets, int tile,
                                                                              int64_t skip_above)
{
    static_assert(SEG_SCAN_PAIRS % WAVE == 0 && SEG_SCAN_PAIRS / WAVE == 2, "two waves of digit pairs");
    __shared__ uint32_t s_part[SEG_SCAN_GROUPS][RADIX];
    __shared__ uint32_t s_wsum[SEG_SCAN_PAIRS / WAVE];
    const int q = threadIdx.x % SEG_SCAN_PAIRS, g = threadIdx.x / SEG_SCAN_PAIRS;
    const int lane = q & (WAVE - 1), wave = q / WAVE;
    // two 16-bit counts per word, two 32-bit offsets per 8 bytes: row t holds
    // SEG_SCAN_PAIRS of each
    const uint32_t *__restrict__ c2 = reinterpret_cast<const uint32_t *>(counts);
    uint2 *__restrict__ o2 = reinterpret_cast<uint2 *>(offsets);
    const uint32_t nlarge = sp->num_large;
    for (uint32_t k = blockIdx.x; k < nlarge; k += gridDim.x) {
        const uint32_t s = large_list[k];
        if (begin[s + 1] - begin[s] > skip_above) continue; // an MSD-first segment after its one pass
        const uint32_t t0 = seg_tile0[s];
        const uint32_t nt = uint32_t((begin[s + 1] - begin[s] + tile - 1) / tile);
        const uint32_t per = (nt + SEG_SCAN_GROUPS - 1) / SEG_SCAN_GROUPS;
        const uint32_t ta = min(nt, uint32_t(g) * per), tz = min(nt, ta + per); // this part's tiles
        uint32_t p0 = 0, p1 = 0;
#pragma unroll 8
        for (uint32_t t = ta; t < tz; ++t) {
            const uint32_t w = c2[size_t(t0 + t) * SEG_SCAN_PAIRS + q];
            p0 += w & 0xFFFFu;
            p1 += w >> 16;
        }
        s_part[g][2 * q] = p0;
        s_part[g][2 * q + 1] = p1;
        __syncthreads();
        // the two digits' totals over the segment and over the earlier parts
        uint32_t tot0 = 0, tot1 = 0, before0 = 0, before1 = 0;
#pragma unroll
        for (int h = 0; h < SEG_SCAN_GROUPS; ++h) {
            const uint32_t v0 = s_part[h][2 * q], v1 = s_part[h][2 * q + 1];
            tot0 += v0;
            tot1 += v1;
            if (h < g) {
                before0 += v0;
                before1 += v1;
            }
        }
        // exclusive scan over the digits: inclusive scan of the pair sums
        const uint32_t pair = tot0 + tot1;
        uint32_t x = pair;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (g == 0 && lane == WAVE - 1) s_wsum[wave] = x;
        __syncthreads();
        const uint32_t below = x - pair + (wave ? s_wsum[0] : 0u); // digits before 2q
        uint32_t run0 = below + before0, run1 = below + tot0 + before1;
        // offsets are positions within the segment (added to its first key at use)
        for (uint32_t t = ta; t < tz; ++t) {
            const size_t row = size_t(t0 + t) * SEG_SCAN_PAIRS + q;
            const uint32_t w = c2[row];
            o2[row] = make_uint2(run0, run1);
            run0 += w & 0xFFFFu;
            run1 += w >> 16;
        }
        __syncthreads();
    }
}

// The large segments' scan for segments of any length (k_seg_scan walks a
// segment's tiles in one workgroup, which is slow for segments of many
// thousands of tiles).  With P(t, d) = the count of digit d over the tile
// records before t (records of one segment are consecutive):
//   k_segx_ranges  per range of SEGX_RANGE records: in-range prefixes into
//                  offsets, range sums into rbase
//   k_segx_columns per digit: exclusive scan of the range sums (rbase now
//                  holds P at each range's first record) and the total
//   k_segx_bases   per large segment k (records t0 .. te): its digit totals
//                  P(te, d) - P(t0, d), their exclusive scan dstart[k][d]
//                  (digit starts within the segment) and sbase[k][d] =
//                  dstart[k][d] - P(t0, d)
// so that the record t of segment k writes digit d from position
// sbase[k][d] + rbase[t / SEGX_RANGE][d] + offsets[t][d] of its segment
// (unsigned arithmetic; the sum lies within the segment).
constexpr int SEGX_RANGE = 64;

static __global__ __launch_bounds__(RADIX) void k_segx_ranges(const SegPlan *__restrict__ sp,
                                                              const uint16_t *__restrict__ counts,
                                                              uint32_t *__restrict__ offsets,
                                                              uint32_t *__restrict__ rbase)
{
    const uint32_t nt = sp->num_tiles;
    const uint32_t r = blockIdx.x, d = threadIdx.x;
    const uint32_t t0 = r * SEGX_RANGE;
    if (t0 >= nt) return;
    const uint32_t t1 = min(nt, t0 + SEGX_RANGE);
    uint32_t c[SEGX_RANGE];
#pragma unroll
    for (int k = 0; k < SEGX_RANGE; ++k) c[k] = t0 + k < t1 ? uint32_t(counts[size_t(t0 + k) * RADIX + d]) : 0u;
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SEGX_RANGE; ++k) {
        if (t0 + k < t1) offsets[size_t(t0 + k) * RADIX + d] = sum;
        sum += c[k];
    }
    rbase[size_t(r) * RADIX + d] = sum;
}

// One workgroup of 1024 threads per digit d.
static __global__ __launch_bounds__(1024) void k_segx_columns(const SegPlan *__restrict__ sp,
                                                              uint32_t *__restrict__ rbase,
                                                              uint32_t *__restrict__ total)
{
    __shared__ uint32_t s_w[1024 / WAVE];
    const uint32_t d = blockIdx.x, tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const uint32_t nr = (sp->num_tiles + SEGX_RANGE - 1) / SEGX_RANGE;
    const uint32_t per = (nr + 1023) / 1024;
    const uint32_t r0 = min(nr, tid * per), r1 = min(nr, r0 + per);
    uint32_t local = 0;
    for (uint32_t r = r0; r < r1; ++r) local += rbase[size_t(r) * RADIX + d];
    uint32_t x = local;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= uint32_t(off)) x += y;
    }
    if (lane == WAVE - 1) s_w[wave] = x;
    __syncthreads();
    uint32_t run = x - local;
    for (uint32_t w = 0; w < wave; ++w) run += s_w[w];
    for (uint32_t r = r0; r < r1; ++r) {
        const size_t at = size_t(r) * RADIX + d;
        const uint32_t c = rbase[at];
        rbase[at] = run;
        run += c;
    }
    if (tid == 1023) total[d] = run; // every range summed (the last thread's run ends at nr)
}

// Grid-stride over the large segments, RADIX threads (one per digit).
static __global__ __launch_bounds__(RADIX) void k_segx_bases(const int64_t *__restrict__ begin,
                                                             const SegPlan *__restrict__ sp,
                                                             const uint32_t *__restrict__ seg_tile0,
                                                             const uint32_t *__restrict__ large_list,
                                                             const uint32_t *__restrict__ offsets,
                                                             const uint32_t *__restrict__ rbase,
                                                             const uint32_t *__restrict__ total,
                                                             uint32_t *__restrict__ sbase, uint32_t *__restrict__ dstart,
                                                             int tile)
{
    __shared__ uint32_t s_w[RADIX / WAVE];
    const uint32_t d = threadIdx.x, lane = d & (WAVE - 1), wave = d / WAVE;
    const uint32_t nlarge = sp->num_large, nt_all = sp->num_tiles;
    auto prefix = [&](uint32_t t) -> uint32_t {
        return t < nt_all ? rbase[size_t(t / SEGX_RANGE) * RADIX + d] + offsets[size_t(t) * RADIX + d] : total[d];
    };
    for (uint32_t k = blockIdx.x; k < nlarge; k += gridDim.x) {
        const uint32_t s = large_list[k];
        const uint32_t t0 = seg_tile0[s];
        const uint32_t nt = uint32_t((begin[s + 1] - begin[s] + tile - 1) / tile);
        const uint32_t p0 = prefix(t0);
        const uint32_t tot = prefix(t0 + nt) - p0;
        uint32_t x = tot;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= uint32_t(off)) x += y;
        }
        if (lane == WAVE - 1) s_w[wave] = x;
        __syncthreads();
        uint32_t start = x - tot;
        for (uint32_t w = 0; w < wave; ++w) start += s_w[w];
        sbase[size_t(k) * RADIX + d] = start - p0;
        if (dstart) dstart[size_t(k) * RADIX + d] = start;
        __syncthreads();
    }
}

// Scatter of pass `pass` for the large segments' tiles, one workgroup per
// tile (spare workgroups of the grid return; STRIDE, as for k_seg_counts:
// a smaller grid walks the tiles grid-stride): the plain sort's prescanned
// scatter (k_onesweep with precomputed offsets) on a tile record.  Workgroup
// b < num_tiles takes tile xcd_contiguous(b, num_tiles): XCD b % 8 walks a
// contiguous share of the tiles, so neighbouring tiles (adjacent digit runs)
// meet in one L2.  The keys are loaded before the LDS set-up; a full wave
// (every wave of a full tile) loads and a full tile ranks and writes without
// per-key bounds tests; RANK ranks as in k_onesweep (RANK_LDSOR counts with
// LDS atomics first and stages while it ranks, RANK_MIX2 / RANK_MIX4 match
// every second / fourth item by ballots, RANK_BALLOT ranks by ballots with
// wave-local counters and stages after the scan).
template <int BLOCK, int ITEMS, int RANK, bool PEEL>
__device__ __forceinline__ void seg_scatter_tile(uint32_t vb, uint32_t ntiles, uint32_t *__restrict__ keysA,
                                                 uint32_t *__restrict__ keysB, int pass,
                                                 const uint4 *__restrict__ tinfo,
                                                 const uint32_t *__restrict__ offsets,
                                                 const uint32_t *__restrict__ rbase,
                                                 const uint32_t *__restrict__ sbase,
                                                 const uint32_t *__restrict__ dshift)
{
    static_assert(BLOCK >= RADIX && BLOCK % WAVE == 0 && BLOCK <= 512, "one thread per digit, at most 8 waves");
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int SCAN_WAVES = RADIX / WAVE;
    __shared__ uint32_t s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[RANK != RANK_BALLOT ? WAVES : 1][RANK != RANK_BALLOT ? RADIX : 1];
    __shared__ uint32_t s_scan[SCAN_WAVES];
    __shared__ uint32_t s_base[RADIX];
    __shared__ uint32_t s_keys[TILE];
    const uint32_t *__restrict__ kin = (pass & 1) ? keysB : keysA;
    uint32_t *__restrict__ kout = (pass & 1) ? keysA : keysB;
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    int shift = dshift ? int(*dshift) : 8 * pass; // the digit's bit; pass picks the buffers
    if (shift >= 32) return;                      // a pass the plan skips
    const uint32_t t = xcd_contiguous(vb, ntiles);
    const uint4 ti = tinfo[t];
    if (ti.w & SEG_MSD_TILE) { // MSD-first segment: pass 0 by the top byte only
        if (pass != 0) return;
        shift = 24;
    }
    const uint32_t tb = ti.x, sb = ti.y;
    const int tile_valid = int(ti.z);
    const int wave_valid = tile_valid > wave * WAVE_TILE ? min(WAVE_TILE, tile_valid - wave * WAVE_TILE) : 0;

    uint32_t keys[ITEMS];
    if (wave_valid == WAVE_TILE) {
        const uint32_t *__restrict__ kp = kin + tb + uint32_t(wave * WAVE_TILE + lane);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) keys[i] = kp[i * WAVE];
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int off = i * WAVE + lane;
            keys[i] = off < wave_valid ? kin[tb + uint32_t(wave * WAVE_TILE + off)] : ~0u;
        }
    }
    // the tile's digit offsets within its segment (k_seg_scan's, or the sum
    // of k_segx_* parts: in-range prefix, range base, segment base)
    uint32_t pre_off = tid < RADIX ? offsets[size_t(t) * RADIX + tid] : 0u;
    if (rbase != nullptr && tid < RADIX)
        pre_off += rbase[size_t(t / SEGX_RANGE) * RADIX + tid] + sbase[size_t(ti.w & ~SEG_MSD_TILE) * RADIX + tid];

    for (int i = tid; i < WAVES * RADIX; i += BLOCK) (&s_cnt[0][0])[i] = 0u;
    if constexpr (RANK != RANK_BALLOT)
        for (int i = tid; i < WAVES * RADIX; i += BLOCK) (&s_match[0][0])[i] = 0ull;
    __syncthreads();

    // ---- per-wave digit counts (ballot ranking: the ranks themselves) ----
    uint32_t *cnt = s_cnt[wave];
    const bool full = tile_valid == TILE;
    const bool skewed = PEEL && RANK != RANK_BALLOT && wave_skewed((keys[0] >> shift) & (RADIX - 1), lane < wave_valid,
                                                           (keys[ITEMS / 2] >> shift) & (RADIX - 1),
                                                           (ITEMS / 2) * WAVE + lane < wave_valid);
    uint32_t rank2[RANK == RANK_BALLOT ? (ITEMS + 1) / 2 : 1];
    if constexpr (RANK == RANK_BALLOT) {
        auto rank_loop = [&](auto full_c) {
            constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const uint32_t d = (keys[i] >> shift) & (RADIX - 1);
                const bool valid = FULL || i * WAVE + lane < wave_valid;
                const uint64_t peers = FULL ? match_digit8(d) : match_digit8(d) & __ballot(valid);
                const uint32_t below = count_below(peers);
                const uint32_t old = cnt[d];
                if (valid && below == 0) cnt[d] = old + uint32_t(__popcll(peers));
                const uint32_t r = old + below;
                rank2[i / 2] = (i & 1) ? (rank2[i / 2] | (r << 16)) : r;
                __builtin_amdgcn_wave_barrier();
            }
        };
        if (full) rank_loop(std::true_type{});
        else rank_loop(std::false_type{});
    } else {
        if (full) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) count_item(cnt, (keys[i] >> shift) & (RADIX - 1), true, lane, skewed);
        } else {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                count_item(cnt, (keys[i] >> shift) & (RADIX - 1), i * WAVE + lane < wave_valid, lane, skewed);
        }
    }
    __syncthreads();

    // ---- per-digit wave bases and the tile's digit starts ----
    uint32_t tile_cnt = 0, tile_start = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = tile_cnt;
            tile_cnt += c;
        }
        uint32_t x = tile_cnt;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == WAVE - 1) s_scan[wave] = x;
        tile_start = x - tile_cnt;
    }
    __syncthreads();
    if (tid < RADIX) {
        for (int w = 0; w < wave; ++w) tile_start += s_scan[w];
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s_cnt[w][tid] += tile_start;
        s_base[tid] = pre_off - tile_start;
    }
    __syncthreads();

    // ---- rank and stage the tile in LDS in sorted order ----
    auto stage_loop = [&](auto full_c, auto peel_c) {
        constexpr bool FULL = decltype(full_c)::value;
        constexpr bool PEELED = decltype(peel_c)::value;
        if constexpr (RANK == RANK_BALLOT) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                if (FULL || i * WAVE + lane < wave_valid)
                    s_keys[cnt[(keys[i] >> shift) & (RADIX - 1)] + ((rank2[i / 2] >> (16 * (i & 1))) & 0xFFFFu)] =
                        keys[i];
        } else {
            uint32_t *match32 = reinterpret_cast<uint32_t *>(s_match[wave]);
            constexpr int MIX = rank_mix_period<RANK>();
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const uint32_t d = (keys[i] >> shift) & (RADIX - 1);
                const bool valid = FULL || i * WAVE + lane < wave_valid;
                if (MIX && i % MIX == MIX - 1) {
                    const uint64_t peers = FULL ? match_digit8(d) : match_digit8(d) & __ballot(valid);
                    const uint32_t below = count_below(peers);
                    const uint32_t old = cnt[d];
                    if (valid && below == 0) cnt[d] = old + uint32_t(__popcll(peers));
                    if (valid) s_keys[old + below] = keys[i];
                    __builtin_amdgcn_wave_barrier();
                } else {
                    const uint32_t at = ldsor_rank_item<PEELED>(d, valid, lane, cnt, match32, s_match[wave]);
                    if (valid) s_keys[at] = keys[i];
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
    };
    if constexpr (RANK == RANK_BALLOT || !PEEL) {
        if (full) stage_loop(std::true_type{}, std::false_type{});
        else stage_loop(std::false_type{}, std::false_type{});
    } else if (full) {
        if (skewed) stage_loop(std::true_type{}, std::true_type{});
        else stage_loop(std::true_type{}, std::false_type{});
    } else {
        if (skewed) stage_loop(std::false_type{}, std::true_type{});
        else stage_loop(std::false_type{}, std::false_type{});
    }
    __syncthreads();

    // ---- write the tile out, each digit's run contiguous ----
    uint32_t *__restrict__ out = kout + sb;
    if (full) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int j = i * BLOCK + tid;
            const uint32_t k = s_keys[j];
            out[s_base[(k >> shift) & (RADIX - 1)] + uint32_t(j)] = k;
        }
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int j = i * BLOCK + tid;
            if (j < tile_valid) {
                const uint32_t k = s_keys[j];
                out[s_base[(k >> shift) & (RADIX - 1)] + uint32_t(j)] = k;
            }
        }
    }
}

template <int BLOCK, int ITEMS, int RANK, bool PEEL = true, bool STRIDE = false>
__global__ __launch_bounds__(BLOCK) void k_seg_scatter(uint32_t *__restrict__ keysA, uint32_t *__restrict__ keysB,
                                                       int pass, const SegPlan *__restrict__ sp,
                                                       const uint4 *__restrict__ tinfo,
                                                       const uint32_t *__restrict__ offsets,
                                                       const uint32_t *__restrict__ rbase,
                                                       const uint32_t *__restrict__ sbase,
                                                       const uint32_t *__restrict__ dshift)
{
    const uint32_t ntiles = sp->num_tiles;
    if constexpr (!STRIDE) {
        if (blockIdx.x < ntiles)
            seg_scatter_tile<BLOCK, ITEMS, RANK, PEEL>(blockIdx.x, ntiles, keysA, keysB, pass, tinfo, offsets, rbase,
                                                       sbase, dshift);
        return;
    }
#pragma nounroll
    for (uint32_t vb = blockIdx.x; vb < ntiles; vb += gridDim.x) {
        seg_scatter_tile<BLOCK, ITEMS, RANK, PEEL>(vb, ntiles, keysA, keysB, pass, tinfo, offsets, rbase, sbase,
                                                   dshift);
        if (vb + gridDim.x < ntiles) __syncthreads(); // the next tile reuses the LDS
    }
}

// ---------------------------------------------------------------------------
// k_sort_block: workgroup b sorts its chunk [b * TILE, min(n, (b + 1) * TILE))
// of keys_in (and u32 values) in LDS into the same place of keys_out (which
// may be keys_in): one read, the LSD passes over the chunk's varying digit
// positions (tile_rank_stage per pass, the float / signed key transform on
// load and its inverse on store), one write.  With one workgroup (n <=
// BLOCK * ITEMS) one launch replaces the count / plan / pass / copy-back
// sequence of the multi-tile sorts; with many, it makes the sorted runs of
// the block-sort + merge sort (hpcs_merge.hpp).
// ---------------------------------------------------------------------------
template <typename K, int FLOATK, bool HAS_VALS, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_sort_block(const K *keys_in, const uint32_t *vals_in, K *keys_out,
                                                      uint32_t *vals_out, uint32_t n_all)
{
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int P = int(sizeof(K));
    // this workgroup's chunk (the grid covers n_all; no workgroup is spare)
    const uint32_t c0 = blockIdx.x * uint32_t(TILE);
    const uint32_t n = min(uint32_t(TILE), n_all - c0);
    const K *keys = keys_in + c0;
    const uint32_t *vals = HAS_VALS ? vals_in + c0 : nullptr;
    // LDS-OR match ranking where its masks fit in LDS beside the keys and
    // values (measured on MI355X against ballot ranking: u32 4096 keys 25.4
    // -> 22.4 us, 16384 48.8 -> 42.4 us, u64 8192 53.5 -> 48.7 us)
    constexpr bool LDSOR = size_t(TILE) * (sizeof(K) + (HAS_VALS ? 4 : 0)) + size_t(WAVES) * RADIX * 12 <=
                           size_t(150) * 1024;
    __shared__ uint32_t s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[LDSOR ? WAVES : 1][LDSOR ? RADIX : 1];
    __shared__ uint32_t s_scan[RADIX / WAVE];
    __shared__ uint32_t s_start[RADIX];
    __shared__ K s_keys[TILE];
    __shared__ uint32_t s_vals[HAS_VALS ? TILE : 1];
    __shared__ K s_or[WAVES], s_nor[WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    const int len = int(min(n, uint32_t(TILE)));
    const int wave_valid = len > wave * WAVE_TILE ? min(WAVE_TILE, len - wave * WAVE_TILE) : 0;
    K v[ITEMS];
    uint32_t pv[HAS_VALS ? ITEMS : 1];
    K vor = 0, vnor = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        const int o = j * WAVE + lane;
        v[j] = ~K(0);
        if (o < wave_valid) {
            v[j] = FLOATK ? key_enc<K, FLOATK>(keys[wave * WAVE_TILE + o]) : keys[wave * WAVE_TILE + o];
            vor |= v[j];
            vnor |= ~v[j];
        }
        if constexpr (HAS_VALS) pv[j] = o < wave_valid ? vals[wave * WAVE_TILE + o] : 0u;
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
    const K varying = a & c; // bits that differ between keys
    bool staged = false;     // keys sit in s_keys in sorted order
    for (int pass = 0; pass < P; ++pass) {
        if (((varying >> (8 * pass)) & K(0xFF)) == 0) continue;
        if (staged) {
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const int o = j * WAVE + lane;
                if (o < wave_valid) {
                    v[j] = s_keys[wave * WAVE_TILE + o];
                    if constexpr (HAS_VALS) pv[j] = s_vals[wave * WAVE_TILE + o];
                }
            }
            __syncthreads();
        }
        if constexpr (LDSOR)
            tile_rank_stage_ldsor<K, BLOCK, ITEMS, HAS_VALS>(v, wave_valid, 8 * pass, s_cnt, s_match, s_scan, s_start,
                                                             s_keys, pv, s_vals);
        else
            tile_rank_stage<K, BLOCK, ITEMS, HAS_VALS>(v, wave_valid, 8 * pass, s_cnt, s_scan, s_start, s_keys, pv,
                                                       s_vals);
        staged = true;
    }
    K *kout = keys_out + c0;
    uint32_t *vout = HAS_VALS ? vals_out + c0 : nullptr;
    if (staged) {
        for (int j = tid; j < len; j += BLOCK) {
            kout[j] = FLOATK ? key_dec<K, FLOATK>(s_keys[j]) : s_keys[j];
            if constexpr (HAS_VALS) vout[j] = s_vals[j];
        }
    } else if (keys_out != keys_in) {
        // every key of the chunk is equal (no pass ran): the chunk as it was
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const int o = j * WAVE + lane;
            if (o < wave_valid) {
                kout[wave * WAVE_TILE + o] = FLOATK ? key_dec<K, FLOATK>(v[j]) : v[j];
                if constexpr (HAS_VALS) vout[wave * WAVE_TILE + o] = pv[j];
            }
        }
    }
}

} // namespace hpcs
