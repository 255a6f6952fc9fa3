// hpcs_digit11.hpp -- the single-pass (look-back) LSD sort with 11-bit
// digits: three passes for 4-byte keys (11 + 11 + 10 bits) and six for 8-byte
// keys (5 x 11 + 9 bits), against four and eight with 8-bit digits.  The
// reference sorts with the same digit split in SortRadixLsd11bit and
// SortRadixLsdNbit (RadixSortLsd.cs:1484-1546, 1616-1752); the result of a
// stable LSD sort does not depend on the digit width.
//
//   k_hist11      every digit's 2048-bin count in one read
//   k_plan11      per pass: digit starts (exclusive scan), trivial passes,
//                 buffer selection, the float/signed key transform's first
//                 and last pass
//   k_onesweep11  one stable scatter pass with decoupled look-back (2048
//                 look-back words per tile)
//   k_copyback11  if the result sits in buffer B, copy it to A
//
// Keys only (no values).  The look-back words carry 30-bit counts, so this
// form serves sorts below 2^30 keys; the library uses it where the tuning
// table or an option asks for 11-bit digits.
#pragma once
#include "hpcs_kernels.hpp"

namespace hpcs {

constexpr int D11_BITS = 11;
constexpr int D11_RADIX = 1 << D11_BITS;
constexpr int D11_MAX_PASSES = 6;
constexpr int D11_BLOCK = 256; // threads per tile; 8 consecutive digits per thread
constexpr int D11_ITEMS = 16;  // keys per thread
constexpr int D11_TILE = D11_BLOCK * D11_ITEMS;
constexpr int D11_WINDOW = 4;  // earlier tiles read per look-back round

template <typename K> constexpr int d11_passes() { return (8 * int(sizeof(K)) + D11_BITS - 1) / D11_BITS; }

template <typename K> __device__ inline uint32_t d11_digit(K k, int pass)
{
    return uint32_t(k >> (pass * D11_BITS)) & uint32_t(D11_RADIX - 1);
}

// Lanes of the wave whose BITS-bit digit equals this lane's (one ballot per
// bit, as match_digit8).
template <int BITS> __device__ inline uint64_t match_bits(uint32_t d)
{
    uint32_t diff_lo = 0u, diff_hi = 0u;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
        const uint32_t sp = uint32_t(__builtin_amdgcn_sbfe(int(d), b, 1));
        const uint64_t m = __ballot(sp != 0u);
        diff_lo |= uint32_t(m) ^ sp;
        diff_hi |= uint32_t(m >> 32) ^ sp;
    }
    return ~((uint64_t(diff_hi) << 32) | diff_lo);
}

struct DevPlan11 {
    uint32_t gstart[D11_MAX_PASSES][D11_RADIX]; // exclusive global start of each digit
    uint32_t flags[D11_MAX_PASSES];             // PF_TRIVIAL | PF_SRC_B | PF_XIN | PF_XOUT
    uint32_t copyback;                          // final data sits in buffer B
    uint32_t pad[57];
};

// Every digit's count, [pass][2048], one read: block tables in LDS (24 KiB
// for 4-byte keys, 48 KiB for 8-byte keys) added into the global table.
template <typename K, int FLOATK>
__global__ __launch_bounds__(256) void k_hist11(const K *__restrict__ keys, uint64_t n, uint32_t *__restrict__ hist)
{
    constexpr int P = d11_passes<K>();
    __shared__ uint32_t sh[P * D11_RADIX];
    for (int i = threadIdx.x; i < P * D11_RADIX; i += 256) sh[i] = 0u;
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t stride = uint64_t(gridDim.x) * 256;
    for (uint64_t base = uint64_t(blockIdx.x) * 256; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < n;
        const K raw = valid ? keys[i] : K(0);
        const K key = FLOATK ? key_enc<K, FLOATK>(raw) : raw;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const uint32_t d = d11_digit(key, p);
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
            if (__all(valid && d == d0)) {
                if (lane == 0) atomicAdd(&sh[p * D11_RADIX + d0], uint32_t(WAVE));
            } else if (valid) {
                atomicAdd(&sh[p * D11_RADIX + d], 1u);
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P * D11_RADIX; i += 256)
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
}

// One block of 1024 threads; thread t owns bins 2t and 2t + 1 of each pass.
template <typename K>
__global__ __launch_bounds__(1024) void k_plan11(const uint32_t *__restrict__ hist, DevPlan11 *__restrict__ plan,
                                                 uint64_t n, int floatk)
{
    constexpr int P = d11_passes<K>();
    __shared__ uint32_t s_w[1024 / WAVE];
    __shared__ uint32_t s_trivial[D11_MAX_PASSES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    if (tid < D11_MAX_PASSES) s_trivial[tid] = 0u;
    __syncthreads();
    for (int p = 0; p < P; ++p) {
        const uint32_t c0 = hist[p * D11_RADIX + 2 * tid], c1 = hist[p * D11_RADIX + 2 * tid + 1];
        if (uint64_t(c0) == n || uint64_t(c1) == n) s_trivial[p] = 1u; // one bin holds every key
        const uint32_t own = c0 + c1;
        uint32_t x = own;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == WAVE - 1) s_w[wave] = x;
        __syncthreads();
        uint32_t excl = x - own;
        for (int w = 0; w < wave; ++w) excl += s_w[w];
        plan->gstart[p][2 * tid] = excl;
        plan->gstart[p][2 * tid + 1] = excl + c0;
        __syncthreads(); // s_w is reused by the next pass
    }
    if (tid == 0) {
        uint32_t cur = 0;
        int first = -1, last = -1;
        uint32_t triv[D11_MAX_PASSES];
        for (int p = 0; p < P; ++p) {
            triv[p] = (n <= 1) ? 1u : s_trivial[p];
            if (!triv[p]) {
                if (first < 0) first = p;
                last = p;
            }
        }
        for (int p = 0; p < P; ++p) {
            uint32_t f = (triv[p] ? PF_TRIVIAL : 0u) | (cur ? PF_SRC_B : 0u);
            if (floatk && p == first) f |= PF_XIN;
            if (floatk && p == last) f |= PF_XOUT;
            plan->flags[p] = f;
            if (!triv[p]) cur ^= 1u;
        }
        plan->copyback = cur;
    }
}

// ---------------------------------------------------------------------------
// k_onesweep11: one stable scatter pass over 11-bit digit `pass`.
//
// Tile = 256 threads x 16 keys; wave w owns the contiguous span
// [w*1024, (w+1)*1024) and loads it lane-striped (item i, lane l -> key
// i*64 + l), so (wave, item, lane) order is input order.  Thread t owns the
// eight consecutive digits [8t, 8t + 8).
// 1. rank: per item, the lanes with the same digit are found with eleven
//    ballots; the group's first lane advances the wave's 16-bit counter of
//    that digit (its final values are the wave's counts).
// 2. per owned digit: per-wave bases, the tile count (published at once as
//    an aggregate), and the tile's digit starts by a block scan.
// 3. each key goes to LDS at its in-tile sorted position.
// 4. decoupled look-back: a round reads D11_WINDOW earlier tiles' words for
//    the thread's eight digits and adds tiles from the newest back while
//    every digit still open is published there, closing a digit at an
//    inclusive count; the tile's inclusive counts are then published.  When
//    the block has waited fb_rounds rounds without progress it counts the
//    digits of the newest unpublished earlier tile itself, so every wait is
//    bounded.
// 5. the tile is written from LDS in order: each digit's run is contiguous.
// ---------------------------------------------------------------------------
template <typename K, int FLOATK>
__global__ __launch_bounds__(D11_BLOCK) void k_onesweep11(K *__restrict__ keysA, K *__restrict__ keysB, uint32_t n,
                                                          int pass, const DevPlan11 *__restrict__ plan,
                                                          uint32_t *__restrict__ lb_cur,
                                                          uint32_t *__restrict__ lb_next, int fb_rounds)
{
    constexpr int BLOCK = D11_BLOCK, ITEMS = D11_ITEMS, TILE = D11_TILE, R = D11_RADIX;
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int DPT = R / BLOCK;                 // digits per thread
    constexpr int W = D11_WINDOW;
    static_assert(DPT * BLOCK == R && DPT <= 32, "digits per thread");
    static_assert(TILE < 65536, "per-wave bases are 16-bit");
    static_assert(WAVES * R * 2 >= R * 4, "the fallback counts reuse the per-wave counters");

    __shared__ uint16_t s_cnt[WAVES][R]; // per-wave counters, then bases; fallback counts in the look-back
    __shared__ uint32_t s_base[R];
    __shared__ uint32_t s_scan[WAVES];
    __shared__ uint32_t s_fb_tile;
    __shared__ K s_keys[TILE];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid >> 6;
    const uint32_t tile = blockIdx.x;
    const int d0 = tid * DPT;

    // clear this tile's look-back words of the next pass (that region's
    // previous pass has finished)
    if (lb_next != nullptr)
#pragma unroll
        for (int k = 0; k < DPT; ++k) lb_next[size_t(tile) * R + d0 + k] = 0u;

    const uint32_t flags = plan->flags[pass];
    if (flags & PF_TRIVIAL) return;
    const K *__restrict__ kin = (flags & PF_SRC_B) ? keysB : keysA;
    K *__restrict__ kout = (flags & PF_SRC_B) ? keysA : keysB;
    const bool xin = FLOATK && (flags & PF_XIN);
    const bool xout = FLOATK && (flags & PF_XOUT);

    const uint32_t tile_base = tile * uint32_t(TILE);
    const uint32_t wave_base = tile_base + uint32_t(wave * WAVE_TILE);
    const int wave_valid = n > wave_base ? int(min(uint32_t(WAVE_TILE), n - wave_base)) : 0;

    K keys[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int off = i * WAVE + lane;
        keys[i] = off < wave_valid ? kin[wave_base + off] : ~K(0);
    }
    for (int i = tid; i < WAVES * R / 2; i += BLOCK) reinterpret_cast<uint32_t *>(&s_cnt[0][0])[i] = 0u;
    __syncthreads();
    if (xin) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            if (i * WAVE + lane < wave_valid) keys[i] = key_enc<K, FLOATK>(keys[i]);
    }

    // ---- 1. rank within the wave ----
    uint16_t *cnt = s_cnt[wave];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t d = d11_digit(keys[i], pass);
        const bool valid = i * WAVE + lane < wave_valid;
        const uint64_t peers = match_bits<D11_BITS>(d) & __ballot(valid);
        const uint32_t below = count_below(peers);
        const uint32_t old = cnt[d];
        if (valid && below == 0) cnt[d] = uint16_t(old + uint32_t(__popcll(peers)));
        rank[i] = old + below;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // ---- 2. wave bases, tile counts (published), tile digit starts ----
    uint32_t tcnt[DPT], tstart[DPT];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w][d0 + k];
            s_cnt[w][d0 + k] = uint16_t(t);
            t += c;
        }
        tcnt[k] = t;
        sum += t;
        lb_store(&lb_cur[size_t(tile) * R + d0 + k], (tile == 0 ? LB_INC : LB_AGG) | t);
    }
    {
        uint32_t x = sum;
#pragma unroll
        for (int off = 1; off < WAVE; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == WAVE - 1) s_scan[wave] = x;
        __syncthreads();
        uint32_t run = x - sum;
        for (int w = 0; w < wave; ++w) run += s_scan[w];
#pragma unroll
        for (int k = 0; k < DPT; ++k) {
            tstart[k] = run;
            run += tcnt[k];
#pragma unroll
            for (int w = 0; w < WAVES; ++w) s_cnt[w][d0 + k] = uint16_t(s_cnt[w][d0 + k] + tstart[k]);
        }
    }
    __syncthreads();

    // ---- 3. stage the tile in LDS in sorted order ----
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
        if (i * WAVE + lane < wave_valid) s_keys[cnt[d11_digit(keys[i], pass)] + rank[i]] = keys[i];

    // ---- 4. decoupled look-back ----
    uint32_t excl[DPT];
#pragma unroll
    for (int k = 0; k < DPT; ++k) excl[k] = 0u;
    uint32_t open = tile == 0 ? 0u : (1u << DPT) - 1u; // digits still summing earlier tiles
    int64_t p = int64_t(tile) - 1;                      // newest earlier tile not yet added
    int idle = 0;                                       // rounds in a row without progress
    uint32_t *s_fb = reinterpret_cast<uint32_t *>(&s_cnt[0][0]);
    while (true) {
        if (open) {
            const int avail = p + 1 < W ? int(p + 1) : W;
            uint32_t v[W][DPT];
#pragma unroll
            for (int t = 0; t < W; ++t)
#pragma unroll
                for (int k = 0; k < DPT; ++k)
                    v[t][k] = (t < avail && ((open >> k) & 1u)) ? lb_load(&lb_cur[size_t(p - t) * R + d0 + k]) : 0u;
            int used = 0;
#pragma unroll
            for (int t = 0; t < W; ++t) {
                if (used == t && t < avail && open) {
                    bool ready = true;
#pragma unroll
                    for (int k = 0; k < DPT; ++k)
                        if (((open >> k) & 1u) && (v[t][k] & ~LB_VALUE) == 0u) ready = false;
                    if (ready) {
                        uint32_t still = open;
#pragma unroll
                        for (int k = 0; k < DPT; ++k) {
                            if ((open >> k) & 1u) {
                                excl[k] += v[t][k] & LB_VALUE;
                                if ((v[t][k] & ~LB_VALUE) == LB_INC) still &= ~(1u << k);
                            }
                        }
                        open = still;
                        used = t + 1;
                    }
                }
            }
            p -= used;
            idle = used ? 0 : idle + 1;
        }
        if (!__syncthreads_or(open != 0u)) break;
        if (__syncthreads_or(open != 0u && idle >= fb_rounds)) {
            // count the newest earlier tile that a waiting thread stopped at
            if (tid == 0) s_fb_tile = 0u;
            for (int i = tid; i < R; i += BLOCK) s_fb[i] = 0u;
            __syncthreads();
            if (open && idle >= fb_rounds && p >= 0) atomicMax(&s_fb_tile, uint32_t(p));
            __syncthreads();
            const uint32_t q = s_fb_tile;
            const uint32_t qbase = q * uint32_t(TILE);
            const uint32_t qn = min(uint32_t(TILE), n - qbase);
            for (uint32_t j = tid; j < qn; j += BLOCK) {
                const K k = kin[qbase + j];
                atomicAdd(&s_fb[d11_digit(xin ? key_enc<K, FLOATK>(k) : k, pass)], 1u);
            }
            __syncthreads();
            if (open && uint32_t(p) == q) {
#pragma unroll
                for (int k = 0; k < DPT; ++k)
                    if ((open >> k) & 1u) excl[k] += s_fb[d0 + k];
                p -= 1;
                idle = 0;
                if (p < 0) open = 0u; // counted every earlier tile
            }
            __syncthreads();
        } else if (open) {
            __builtin_amdgcn_s_sleep(8);
        }
    }
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
        if (tile > 0) lb_store(&lb_cur[size_t(tile) * R + d0 + k], LB_INC | (excl[k] + tcnt[k]));
        s_base[d0 + k] = plan->gstart[pass][d0 + k] + excl[k] - tstart[k];
    }
    __syncthreads();

    // ---- 5. write the tile out ----
    const int tile_valid = n > tile_base ? int(min(uint32_t(TILE), n - tile_base)) : 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int j = i * BLOCK + tid;
        if (j < tile_valid) {
            const K k = s_keys[j];
            kout[s_base[d11_digit(k, pass)] + uint32_t(j)] = xout ? key_dec<K, FLOATK>(k) : k;
        }
    }
}

template <typename K>
__global__ void k_copyback11(K *__restrict__ keysA, const K *__restrict__ keysB, uint64_t n,
                             const DevPlan11 *__restrict__ plan)
{
    if (!plan->copyback) return;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) keysA[i] = keysB[i];
}

} // namespace hpcs
