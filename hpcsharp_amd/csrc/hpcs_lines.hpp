// hpcs_lines.hpp -- the reduce-then-scan scatter with line-carrying chunks.
//
// k_scatter_lines: one stable scatter pass over 8-bit digit `pass` for
// 4-byte keys without values, on the prescanned offsets that k_onesweep's
// reduce-then-scan form reads (offsets[tile][digit] = global position of the
// tile's first key of each digit).  It sorts what that form sorts, with a
// different write pattern.
//
// Workgroup b handles a chunk of `chunk` consecutive tiles (chunks dealt to
// the XCDs contiguously, as xcd_contiguous deals tiles).  Inside a chunk the
// run of digit d in tile t+1 starts where tile t's run ends, so the chunk's
// keys of digit d fill one contiguous span of the output.  The workgroup
// writes that span in whole 128-byte lines:
//   1. a tile is counted, ranked and staged in LDS in digit order exactly as
//      k_onesweep does (LDS atomics for the counts, LDS-OR match ranking);
//   2. thread d plans digit d: written-through position w (everything of the
//      digit before w is in the output), the tile's run [o, e), and the write
//      limit lim = the last line boundary at or below e (e itself on the
//      chunk's last tile); it lists the lines meeting [w, lim) in a line table;
//   3. each group of 32 lanes writes one listed line: positions below o come
//      from the digit's carry (keys of earlier tiles that completed no line,
//      slot = position mod 32), the others from the staged tile;
//   4. the staged keys at or beyond lim go to the carry for the next tile.
// The carry of a digit never spans two lines (it holds fewer than 32 keys
// that all lie in the line of w), so its slots are distinct.  Each output
// line is written by one store instruction of one lane group, except the
// partial lines at chunk ends, instead of runs of about 16 keys that split
// most lines between two workgroups.
//
// The keys of the next tile (and its offsets) are loaded while the current
// tile is ranked and written.
#pragma once
#include "hpcs_kernels.hpp"

namespace hpcs {

template <typename K, int FLOATK, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_scatter_lines(K *__restrict__ keysA, K *__restrict__ keysB, uint32_t n,
                                                         int pass, const DevPlan *__restrict__ plan,
                                                         const uint32_t *__restrict__ offsets, uint32_t tiles,
                                                         uint32_t chunk)
{
    static_assert(BLOCK == RADIX, "thread d plans digit d");
    static_assert(sizeof(K) == 4, "4-byte keys: 32 keys per 128-byte line");
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int LK = 128 / int(sizeof(K)); // keys per line
    constexpr int GROUPS = BLOCK / LK;       // lane groups, one line each per round
    // lines listed per tile, at most: sum over digits of (carry + run) / LK + 2
    constexpr int MAXL = 3 * RADIX + TILE / LK;

    __shared__ uint32_t s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[WAVES][RADIX];
    __shared__ uint32_t s_scan[WAVES];
    __shared__ uint32_t s_lscan[WAVES];
    __shared__ K s_keys[TILE];
    __shared__ K s_carry[RADIX][LK];
    __shared__ uint4 s_info[RADIX]; // per digit: w, o, lim, first staged index
    __shared__ uint2 s_bl[RADIX];   // per digit: o - first staged index, lim
    __shared__ uint32_t s_desc[MAXL];
    __shared__ uint32_t s_nlines;

    const uint32_t flags = plan->flags[pass];
    if (flags & PF_TRIVIAL) return;
    const uint32_t src = flags & PF_SRC_B;
    const K *__restrict__ kin = src ? keysB : keysA;
    K *__restrict__ kout = src ? keysA : keysB;
    const bool xin = FLOATK && (flags & PF_XIN);
    const bool xout = FLOATK && (flags & PF_XOUT);
    const int shift = pass * RADIX_BITS;

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid >> 6;

    const uint32_t t0 = xcd_contiguous(blockIdx.x, gridDim.x) * chunk;
    if (t0 >= tiles) return; // the whole workgroup
    const uint32_t t1 = min(tiles, t0 + chunk);

    auto load_tile = [&](uint32_t t, K (&k)[ITEMS]) {
        const uint32_t wb = t * uint32_t(TILE) + uint32_t(wave * WAVE_TILE);
        const int wv = n > wb ? int(min(uint32_t(WAVE_TILE), n - wb)) : 0;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int off = i * WAVE + lane;
            k[i] = off < wv ? kin[wb + off] : ~K(0);
        }
    };

    K keys[ITEMS];
    load_tile(t0, keys);
    uint32_t off = offsets[size_t(t0) * RADIX + tid]; // o of digit tid in the current tile
    uint32_t wpos = off;                               // written-through position of digit tid

    for (uint32_t t = t0; t < t1; ++t) {
        const bool last = t + 1 == t1;
        K nkeys[ITEMS];
        uint32_t noff = 0;
        if (!last) {
            load_tile(t + 1, nkeys);
            noff = offsets[size_t(t + 1) * RADIX + tid];
        }
        const uint32_t tile_base = t * uint32_t(TILE);
        const uint32_t wave_base = tile_base + uint32_t(wave * WAVE_TILE);
        const int wave_valid = n > wave_base ? int(min(uint32_t(WAVE_TILE), n - wave_base)) : 0;
        const int tile_valid = n > tile_base ? int(min(uint32_t(TILE), n - tile_base)) : 0;

        for (int i = tid; i < WAVES * RADIX; i += BLOCK) {
            (&s_cnt[0][0])[i] = 0u;
            (&s_match[0][0])[i] = 0ull;
        }
        __syncthreads();
        if (xin) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                if (i * WAVE + lane < wave_valid) keys[i] = key_enc<K, FLOATK>(keys[i]);
        }

        // ---- per-wave digit counts ----
        uint32_t *cnt = s_cnt[wave];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            if (i * WAVE + lane < wave_valid) atomicAdd(&cnt[KeyOps<K>::digit(keys[i], shift)], 1u);
        __syncthreads();

        // ---- per-digit wave bases, tile count, scan over digits ----
        uint32_t tile_cnt = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = tile_cnt;
            tile_cnt += c;
        }
        uint32_t x = tile_cnt;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == WAVE - 1) s_scan[wave] = x;
        uint32_t tile_start = x - tile_cnt;
        __syncthreads();
        for (int w = 0; w < wave; ++w) tile_start += s_scan[w];
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s_cnt[w][tid] += tile_start;
        __syncthreads();

        // ---- rank (LDS-OR match) and stage the tile in digit order ----
        {
            uint32_t *match32 = reinterpret_cast<uint32_t *>(s_match[wave]);
            const uint32_t lanebit = 1u << (lane & 31);
            const int half = lane >> 5;
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const uint32_t d = KeyOps<K>::digit(keys[i], shift);
                const bool valid = i * WAVE + lane < wave_valid;
                if (valid) atomicOr(&match32[2 * d + half], lanebit);
                __builtin_amdgcn_wave_barrier();
                const uint64_t peers = s_match[wave][d];
                const uint32_t below = count_below(peers);
                const uint32_t old = cnt[d];
                if (valid && below == 0) {
                    cnt[d] = old + uint32_t(__popcll(peers));
                    s_match[wave][d] = 0ull;
                }
                if (valid) s_keys[old + below] = keys[i];
                __builtin_amdgcn_wave_barrier();
            }
        }

        // ---- line plan of digit tid ----
        const uint32_t e = off + tile_cnt;
        uint32_t lim = last ? e : (e & ~uint32_t(LK - 1));
        if (lim < wpos) lim = wpos;
        const uint32_t nl = lim > wpos ? (lim - 1) / LK - wpos / LK + 1 : 0u;
        s_info[tid] = make_uint4(wpos, off, lim, tile_start);
        s_bl[tid] = make_uint2(off - tile_start, lim);
        uint32_t ls = nl;
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t y = __shfl_up(ls, o);
            if (lane >= o) ls += y;
        }
        if (lane == WAVE - 1) s_lscan[wave] = ls;
        __syncthreads(); // also: the staged tile is complete
        uint32_t lp = ls - nl;
        for (int w = 0; w < wave; ++w) lp += s_lscan[w];
        for (uint32_t k = 0; k < nl; ++k) s_desc[lp + k] = (uint32_t(tid) << 24) | k;
        if (tid == BLOCK - 1) s_nlines = lp + nl;
        __syncthreads();

        // ---- write the listed lines, one per lane group ----
        {
            const uint32_t nlines = s_nlines;
            const int g = tid / LK, gl = tid % LK;
            for (uint32_t li = uint32_t(g); li < nlines; li += GROUPS) {
                const uint32_t dsc = s_desc[li];
                const uint32_t d = dsc >> 24;
                const uint4 inf = s_info[d];
                const uint32_t q = (inf.x / LK + (dsc & 0xFFFFFFu)) * LK + uint32_t(gl);
                if (q >= inf.x && q < inf.z) {
                    const K k = q < inf.y ? s_carry[d][q % LK] : s_keys[inf.w + (q - inf.y)];
                    kout[q] = xout ? key_dec<K, FLOATK>(k) : k;
                }
            }
        }
        __syncthreads(); // the carry slots just read may be refilled below

        // ---- keys that complete no line stay for the next tile ----
        if (!last) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const int j = i * BLOCK + tid;
                if (j < tile_valid) {
                    const K k = s_keys[j];
                    const uint32_t d = KeyOps<K>::digit(k, shift);
                    const uint2 bl = s_bl[d];
                    const uint32_t q = bl.x + uint32_t(j);
                    if (q >= bl.y) s_carry[d][q % LK] = k;
                }
            }
        }
        wpos = lim;
        off = noff;
        if (!last) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) keys[i] = nkeys[i];
        }
        __syncthreads(); // LDS is reused by the next tile
    }
}

} // namespace hpcs
