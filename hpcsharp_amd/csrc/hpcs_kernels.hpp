// hpcs_kernels.hpp -- gfx950 kernels of the stable LSD radix sort.
//
// Two pipelines sort n keys (8-bit digits, P = sizeof(key) passes).
//
// Reduce-then-scan (large arrays): k_upsweep (pass-0 tile counts and the
// OR / AND of all keys in one read), k_plan_rts (which passes move data),
// then per pass k_tile_counts (passes >= 1), k_range_totals /
// k_scan_columns (every tile's digit offsets) and
// k_onesweep in its prescanned form as the scatter (tiles XCD-contiguous).
//
// Onesweep (small arrays):
//   k_fill        zero the histogram, the plan and look-back region 0
//   k_histogram   one read of the keys, every digit's 256-bin count at once
//                 (the GPU form of HistogramByteComponents, Histogram.cs:261-296)
//   k_plan        per pass: exclusive digit starts (ComputeStartOfBins,
//                 RadixSortLsd.cs:112-125), whether the pass is trivial (all
//                 keys share the digit: a stable identity), ping-pong buffer
//                 selection, where the float key transform goes in and out
//   k_onesweep    one scatter pass (SortRadixPar's per-quantum pass,
//                 RadixSortLsdParallel.cs:184-226): tiles take ids in launch
//                 order, rank their keys stably in LDS, find their global
//                 offsets by decoupled look-back over earlier tiles, and write
//                 each digit's run contiguously from LDS
//   k_copyback    if an odd number of passes ran, move the result back
//
// Stability: the tile order equals the input order and the in-tile rank
// follows (wave, item, lane) = input order, so equal digits keep their order
// in every pass, which is what makes LSD radix sort correct.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace hpcs {

constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int MAX_PASSES = 8;
constexpr int WAVE = 64;

// Look-back word: [31:30] status, [29:0] count.  Status 0 = not yet
// published, AGG = this tile's own count, INC = count of all tiles up to and
// including this one.
constexpr uint32_t LB_AGG = 1u << 30;
constexpr uint32_t LB_INC = 2u << 30;
constexpr uint32_t LB_VALUE = (1u << 30) - 1;
// A look-back word holds counts below 2^30, so the onesweep pass serves
// fewer keys; every other path takes up to the reference's largest array
// (2^31 - 1 elements) with 32-bit counts and offsets.
constexpr int64_t LOOKBACK_MAX_KEYS = int64_t(1) << 30;
constexpr int64_t MAX_KEYS = int64_t(1) << 31;

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint64_t u64x2 __attribute__((ext_vector_type(2)));
template <typename K> struct Vec16;
template <> struct Vec16<uint32_t> { using type = u32x4; };
template <> struct Vec16<uint64_t> { using type = u64x2; };

// Device-side plan, written by k_plan and read by the pass kernels.  It sits in
// the zero-filled part of the scratch so that the tile counters start at 0.
struct DevPlan {
    uint32_t gstart[MAX_PASSES][RADIX]; // exclusive global start of each digit
    uint32_t trivial[MAX_PASSES];       // 1 if every key has the same digit
    uint32_t src[MAX_PASSES];           // 0: data in buffer A, 1: in buffer B
    uint32_t xin[MAX_PASSES];           // apply the float key transform on load
    uint32_t xout[MAX_PASSES];          // apply the inverse transform on store
    uint32_t copyback;                  // final data sits in buffer B
    uint32_t flags[MAX_PASSES];         // the four fields above packed, PF_*
    // MSD split (k_plan_msd): one scatter by the 8-bit digit at msd_shift,
    // the top eight bits in which the keys differ
    uint32_t msd_flags;                 // PF_* of that scatter (PF_TRIVIAL: every key equal)
    uint32_t msd_shift;
    uint32_t msd_start[RADIX];          // digit counts (k_msd_hist), then exclusive starts
    // MSD-first form of the 4-byte keys-only sort (k_plan_rts with `hybrid`):
    // the stages' mode (HYB_*), the digit positions of its global pass slot
    // (pass 3 of the plan) and of its segmented pass, and the digit
    // positions of the two overflow passes (HYB_SKIP: the pass does nothing)
    uint32_t hybrid;
    uint32_t hyb_shift1, hyb_shift2;
    uint32_t hyb_ov_shift[2];
    uint32_t pad[40];
};
constexpr uint32_t PF_TRIVIAL = 1u, PF_SRC_B = 2u, PF_XIN = 4u, PF_XOUT = 8u;
// The pass's tile counts were written by the upsweep (the MSD-first form's
// top-byte pass): its counting kernel does not run, its range kernel reads
// the upsweep's counts.
constexpr uint32_t PF_PRECOUNTED = 32u;
// The pass's digit sits at bit (flags >> PF_SHIFT_POS) & 63 of the key, not
// at 8 * pass (the MSD-first form's first split).
constexpr uint32_t PF_SHIFT = 64u;
constexpr int PF_SHIFT_POS = 16;
// The pass's tile counts are the upsweep's byte-0 counts, already in the
// counts array: its counting kernel does not run.
constexpr uint32_t PF_PRE0 = 128u;
// Modes of the MSD-first form's stages (DevPlan::hybrid): nothing to do;
// the MSD-first form (top varying bit 23 or higher); the LSD passes over the
// varying bytes of keys below 2^23, one per stage slot.
constexpr uint32_t HYB_NONE = 0u, HYB_MSD = 1u, HYB_LSD = 2u;
constexpr uint32_t HYB_SKIP = 32u; // a digit position meaning "no such pass"
__device__ inline int pass_shift(uint32_t flags, int pass)
{
    return (flags & PF_SHIFT) ? int((flags >> PF_SHIFT_POS) & 63u) : pass * RADIX_BITS;
}
// Reduce-then-scan plans only: PF_FIRST marks the first pass that moves keys,
// and bits PF_NEXT_SHIFT.. of a pass's flags hold the next pass that moves
// keys (PF_NEXT_NONE after the last one).
constexpr uint32_t PF_FIRST = 16u;
constexpr int PF_NEXT_SHIFT = 8;
constexpr uint32_t PF_NEXT_NONE = 0xFu;

// Pack the per-pass fields into flags[] (one thread, after they are set).
__device__ inline void pack_plan_flags(DevPlan *plan, int P)
{
    for (int p = 0; p < P; ++p)
        plan->flags[p] = (plan->trivial[p] ? PF_TRIVIAL : 0u) | (plan->src[p] ? PF_SRC_B : 0u) |
                         (plan->xin[p] ? PF_XIN : 0u) | (plan->xout[p] ? PF_XOUT : 0u);
}

// ---------------------------------------------------------------------------
// Key traits: the float/double radix key of RadixSortMsd.cs:1379-1383 and
// :1484-1488 (non-negative: flip the sign bit; negative: flip every bit).
// ---------------------------------------------------------------------------
template <typename K> struct KeyOps;
template <> struct KeyOps<uint32_t> {
    static constexpr uint32_t SIGN = 0x80000000u;
    __device__ static inline uint32_t to_key(uint32_t b) { return (b & SIGN) ? ~b : (b ^ SIGN); }
    __device__ static inline uint32_t from_key(uint32_t k) { return (k & SIGN) ? (k ^ SIGN) : ~k; }
    __device__ static inline uint32_t digit(uint32_t k, int shift) { return (k >> shift) & (RADIX - 1); }
};
template <> struct KeyOps<uint64_t> {
    static constexpr uint64_t SIGN = 0x8000000000000000ull;
    __device__ static inline uint64_t to_key(uint64_t b) { return (b & SIGN) ? ~b : (b ^ SIGN); }
    __device__ static inline uint64_t from_key(uint64_t k) { return (k & SIGN) ? (k ^ SIGN) : ~k; }
    __device__ static inline uint32_t digit(uint64_t k, int shift) { return uint32_t(k >> shift) & (RADIX - 1); }
};

// Key kinds: 0 unsigned (the bits are the key), 1 IEEE float (sign flip for
// non-negatives, all bits flipped for negatives: Histogram.cs:1210-1288,
// RadixSortMsd.cs:1379-1383), 2 two's-complement signed (sign bit flipped:
// the reference's top digit ^ 128, RadixSortLsd.cs:293, Histogram.cs:1019).
constexpr int KIND_UINT = 0, KIND_FLOAT = 1, KIND_INT = 2;
template <typename K, int KIND> __device__ inline K key_enc(K b)
{
    return KIND == KIND_FLOAT ? KeyOps<K>::to_key(b) : KIND == KIND_INT ? K(b ^ KeyOps<K>::SIGN) : b;
}
template <typename K, int KIND> __device__ inline K key_dec(K k)
{
    return KIND == KIND_FLOAT ? KeyOps<K>::from_key(k) : KIND == KIND_INT ? K(k ^ KeyOps<K>::SIGN) : k;
}

__device__ inline uint32_t lb_load(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void lb_store(uint32_t *p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lanes of the wave whose 8-bit digit equals this lane's: eight ballots, one
// per digit bit.  Lane j differs from this lane in bit b iff ballot_b[j]
// differs from this lane's bit, i.e. (ballot_b ^ splat_b)[j] = 1 with
// splat_b = all-ones when this lane's bit is set; the peers are the lanes
// that differ in no bit.
__device__ inline uint64_t match_digit8(uint32_t d)
{
    uint32_t diff_lo = 0u, diff_hi = 0u;
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b) {
        const uint32_t sp = uint32_t(__builtin_amdgcn_sbfe(int(d), b, 1)); // ~0 if bit b set
        const uint64_t m = __ballot(sp != 0u);
        diff_lo |= uint32_t(m) ^ sp;         // one three-input bit operation per half
        diff_hi |= uint32_t(m >> 32) ^ sp;
    }
    return ~((uint64_t(diff_hi) << 32) | diff_lo);
}

// Workgroup b runs on XCD b % 8; map it to tile start(b % 8) + b / 8 so that
// XCD x processes the contiguous tiles [start(x), start(x + 1)).  A bijection
// on [0, g) for any g.
constexpr uint32_t NUM_XCD = 8;
__device__ inline uint32_t xcd_contiguous(uint32_t b, uint32_t g)
{
    const uint32_t x = b % NUM_XCD, i = b / NUM_XCD;
    const uint32_t per = g / NUM_XCD, rem = g % NUM_XCD;
    return x * per + min(x, rem) + i;
}

// Number of set bits of `mask` below this lane.
__device__ inline uint32_t count_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// The valid lanes that hold the first lane's digit, when at least PEEL_MIN
// do (else 0; a wave-uniform value).  Many lanes of a wave with one digit
// (sorted or skewed keys) would queue their LDS atomics on one address; these
// lanes are counted and ranked by this one ballot instead.
constexpr int PEEL_MIN = 16;
__device__ inline uint64_t peel_mask(uint32_t d, bool valid)
{
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const uint64_t m0 = __ballot(valid && d == d0);
    return __popcll(m0) >= PEEL_MIN ? m0 : 0ull;
}

// cnt[d] += 1 for every valid lane, one LDS atomic per lane outside the
// peeled group and one for the group.
template <typename C>
__device__ inline void count_peeled(C *cnt, uint32_t d, bool valid, int lane)
{
    const uint64_t m0 = peel_mask(d, valid);
    if (m0) {
        if (lane == int(__builtin_ctzll(m0))) atomicAdd(&cnt[d], uint32_t(__popcll(m0)));
        else if (valid && !((m0 >> lane) & 1ull)) atomicAdd(&cnt[d], 1u);
    } else if (valid) {
        atomicAdd(&cnt[d], 1u);
    }
}

// Whether a wave's keys look skewed: many lanes share the first lane's digit
// in the first or the middle item (sorted or dominated inputs).  Such waves
// count and rank with peeling (count_peeled, ldsor_rank_item<true>); the
// others skip its per-item test.
__device__ inline bool wave_skewed(uint32_t da, bool va, uint32_t db, bool vb)
{
    return (peel_mask(da, va) | peel_mask(db, vb)) != 0ull;
}

// cnt[d] += 1 for every valid lane: count_peeled in skewed waves, else one
// LDS atomic per lane.
template <typename C>
__device__ inline void count_item(C *cnt, uint32_t d, bool valid, int lane, bool skewed)
{
    if (skewed) count_peeled(cnt, d, valid, lane);
    else if (valid) atomicAdd(&cnt[d], 1u);
}

// One item of LDS-OR ranking in wave-local tables: each valid lane ORs its
// bit into its digit's 64-bit mask (match, as 32-bit halves in match32),
// reads the mask back (its peers), and the lowest peer advances the wave's
// counter cnt[d] by the peer count and clears the mask.  With PEEL a peeled
// group takes its mask from the ballot.  Returns the lane's rank: cnt[d]
// before the item plus the peers below the lane (lane order, so stable).
template <bool PEEL, typename C>
__device__ inline uint32_t ldsor_rank_item(uint32_t d, bool valid, int lane, C *cnt, uint32_t *match32,
                                           uint64_t *match)
{
    const uint32_t lanebit = 1u << (lane & 31);
    const int half = lane >> 5;
    const uint64_t m0 = PEEL ? peel_mask(d, valid) : 0ull;
    uint64_t peers;
    bool mine0 = false;
    if (PEEL && m0) {
        mine0 = (m0 >> lane) & 1ull;
        if (valid && !mine0) atomicOr(&match32[2 * d + half], lanebit);
        __builtin_amdgcn_wave_barrier();
        peers = mine0 ? m0 : match[d];
    } else {
        if (valid) atomicOr(&match32[2 * d + half], lanebit);
        __builtin_amdgcn_wave_barrier();
        peers = match[d];
    }
    const uint32_t below = count_below(peers);
    const uint32_t old = cnt[d];
    if (valid && below == 0) {
        cnt[d] = C(old + uint32_t(__popcll(peers)));
        if (!mine0) match[d] = 0ull;
    }
    return old + below;
}

// ---------------------------------------------------------------------------
// k_fill: zero (or fill) a word range.  Used instead of hipMemsetAsync so the
// sequence stays capturable in a HIP graph.
// ---------------------------------------------------------------------------
static __global__ void k_fill_u32(uint32_t *__restrict__ p, uint64_t words, uint32_t value)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = value;
}

// ---------------------------------------------------------------------------
// k_histogram: every digit's 256-bin count in one read.
// Waves walk 1 KiB chunks (64 lanes x one 16-byte vector) in a grid-stride
// loop; each digit update is an LDS atomic on the block's table, except that
// a chunk whose keys all share a digit adds its key count once (sorted and
// constant inputs).  That test runs once per digit and chunk, on the OR over
// a lane's keys of key XOR lane 0's first key; a test per key and digit cost
// more instructions than the atomics it saved.  The table holds COPIES
// interleaved copies of the [P][256] counts (32-bit), word (d * 256 + bin) *
// COPIES + lane % COPIES, so that a wave's atomics on random digits collide
// less on LDS banks: with 16 copies (4-byte keys) or 8 (8-byte keys) the
// table fills 64 KiB, two workgroups per CU.  UNROLL chunks per wave are in
// flight.  The tail below a whole chunk is counted lane by lane.  Block
// tables (copies summed) are added into the global [P][256] result with one
// atomic per non-zero bin.
// ---------------------------------------------------------------------------
// `head` (< VEC) keys precede the first 16-byte aligned key; they are counted
// lane by lane together with the tail.

// lane 0's value of x, for every lane
__device__ inline uint32_t wave_first(uint32_t x) { return uint32_t(__builtin_amdgcn_readfirstlane(int(x))); }
__device__ inline uint64_t wave_first(uint64_t x)
{
    return uint64_t(wave_first(uint32_t(x))) | (uint64_t(wave_first(uint32_t(x >> 32))) << 32);
}
template <typename K, int FLOATK, int BLOCK, int COPIES, int UNROLL>
__global__ __launch_bounds__(BLOCK) void k_histogram(const K *__restrict__ keys, uint64_t n, uint32_t head,
                                                     uint32_t *__restrict__ hist)
{
    constexpr int P = int(sizeof(K));
    constexpr int VEC = 16 / int(sizeof(K)); // keys per 16-byte vector
    constexpr int CHUNK = WAVE * VEC;
    constexpr int WAVES = BLOCK / WAVE;
    using Vec = typename Vec16<K>::type;
    __shared__ uint32_t sh[P * RADIX * COPIES];

    for (int i = threadIdx.x; i < P * RADIX * COPIES; i += BLOCK) sh[i] = 0;
    __syncthreads();

    const int lane = threadIdx.x & (WAVE - 1);
    const int copy = lane % COPIES;
    const uint64_t body = n > head ? n - head : 0;
    const uint64_t nchunks = body / CHUNK;
    const uint64_t wave_id = uint64_t(blockIdx.x) * WAVES + (threadIdx.x >> 6);
    const uint64_t nwaves = uint64_t(gridDim.x) * WAVES;
    const Vec *vkeys = reinterpret_cast<const Vec *>(keys + head);

    // UNROLL chunks per wave and round, all loads issued before any is
    // counted (one 1 KiB load per wave in flight left HBM mostly idle)
    for (uint64_t c0 = wave_id * UNROLL; c0 < nchunks; c0 += nwaves * UNROLL) {
        Vec v[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
            if (c0 + u < nchunks) v[u] = __builtin_nontemporal_load(&vkeys[(c0 + u) * WAVE + lane]);
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            if (c0 + u >= nchunks) break; // wave-uniform
            K kk[VEC];
#pragma unroll
            for (int j = 0; j < VEC; ++j) kk[j] = FLOATK ? key_enc<K, FLOATK>(v[u][j]) : v[u][j];
            // the bits in which some key of the chunk differs from lane 0's
            // first key: digit d is the same for the whole chunk iff its
            // byte of `m` is zero in every lane (one test per digit and chunk)
            const K ref = wave_first(kk[0]);
            K m = 0;
#pragma unroll
            for (int j = 0; j < VEC; ++j) m |= kk[j] ^ ref;
#pragma unroll
            for (int d = 0; d < P; ++d) {
                if (__all(KeyOps<K>::digit(m, 8 * d) == 0u)) {
                    if (lane == 0) atomicAdd(&sh[(d * RADIX + KeyOps<K>::digit(ref, 8 * d)) * COPIES], uint32_t(CHUNK));
                } else {
#pragma unroll
                    for (int j = 0; j < VEC; ++j)
                        atomicAdd(&sh[(d * RADIX + KeyOps<K>::digit(kk[j], 8 * d)) * COPIES + copy], 1u);
                }
            }
        }
    }
    // head keys [0, head) and tail keys [head + nchunks * CHUNK, n), block 0
    if (blockIdx.x == 0) {
        const uint64_t tail0 = head + nchunks * CHUNK;
        const uint64_t extra = (n < head ? n : uint64_t(head)) + (n > tail0 ? n - tail0 : 0);
        for (uint64_t t = threadIdx.x; t < extra; t += BLOCK) {
            const uint64_t i = t < head ? t : tail0 + (t - head);
            K key = FLOATK ? key_enc<K, FLOATK>(keys[i]) : keys[i];
#pragma unroll
            for (int d = 0; d < P; ++d) atomicAdd(&sh[(d * RADIX + KeyOps<K>::digit(key, 8 * d)) * COPIES + copy], 1u);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) {
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < COPIES; ++k) c += sh[i * COPIES + k];
        if (c) atomicAdd(&hist[i], c);
    }
}

// ---------------------------------------------------------------------------
// k_plan: one block of RADIX threads; thread b owns bin b of every pass.
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(RADIX) void k_plan(const uint32_t *__restrict__ hist, DevPlan *__restrict__ plan,
                                                uint64_t n, int floatk)
{
    __shared__ uint32_t s[RADIX];
    __shared__ uint32_t s_trivial[MAX_PASSES];
    const int b = threadIdx.x;
    if (b < MAX_PASSES) s_trivial[b] = 0;
    __syncthreads();
    for (int p = 0; p < P; ++p) {
        const uint32_t c = hist[p * RADIX + b];
        if (uint64_t(c) == n) s_trivial[p] = 1; // one bin holds every key
        s[b] = c;
        __syncthreads();
        // Hillis-Steele inclusive scan over 256 bins
        for (int off = 1; off < RADIX; off <<= 1) {
            const uint32_t add = b >= off ? s[b - off] : 0u;
            __syncthreads();
            s[b] += add;
            __syncthreads();
        }
        plan->gstart[p][b] = s[b] - c;
        __syncthreads();
    }
    if (b == 0) {
        uint32_t cur = 0;
        int first = -1, last = -1;
        for (int p = 0; p < P; ++p) {
            const uint32_t t = (n <= 1) ? 1u : s_trivial[p];
            plan->trivial[p] = t;
            plan->src[p] = cur;
            if (!t) {
                cur ^= 1u;
                if (first < 0) first = p;
                last = p;
            }
        }
        for (int p = 0; p < P; ++p) {
            plan->xin[p] = (floatk && p == first) ? 1u : 0u;
            plan->xout[p] = (floatk && p == last) ? 1u : 0u;
        }
        plan->copyback = cur;
        pack_plan_flags(plan, P);
    }
}

// ---------------------------------------------------------------------------
// k_onesweep: one stable scatter pass over 8-bit digit `pass`.
//
// Tile = BLOCK threads x ITEMS keys.  Wave w of the tile owns the contiguous
// span [w*64*ITEMS, (w+1)*64*ITEMS) and loads it lane-striped (item i, lane l
// -> key i*64 + l), so every load instruction reads 64 consecutive keys and
// (wave, item, lane) order is input order.
//
// 1. count: each wave counts its digits with LDS atomics (order-free).
// 2. thread b (one per digit) turns the per-wave counts into per-wave bases,
//    publishes the tile's own count for the look-back at once (so later
//    tiles rarely wait on it), and scans the tile counts over digits.
// 3. rank + stage: items in order; the lanes holding the same digit are found
//    (RANK), the group's first lane advances the wave's slot counter for that
//    digit, and each key goes to LDS at its in-tile sorted position.
// 4. decoupled look-back: thread b reads LOOKBACK_WINDOW earlier tiles' words
//    for digit b in one round trip, adds own-counts from the newest back until
//    it meets an inclusive count, and repeats only if the window ran out or
//    met an unpublished tile; then it publishes its own inclusive count.
//    The tile id is blockIdx.x (no global counter: a single-address atomic
//    per tile serialises at the memory side).  Nothing assumes that earlier
//    tiles are already running: when the block has waited FALLBACK_ROUNDS
//    rounds on an unpublished tile, it reads that tile's keys and counts its
//    digits itself (decoupled fallback), so every wait is bounded.
// 5. the tile is written from LDS in order; each digit's run lands at
//    consecutive addresses.
//
// RANK selects how a wave finds the lanes that share a digit:
//   RANK_BALLOT  eight ballots (VALU work, no LDS traffic);
//   RANK_LDSOR   each lane ORs its lane bit into a per-wave LDS mask of its
//                digit and reads the mask back (LDS work, few VALU ops); OR is
//                commutative, so the mask does not depend on the order in
//                which the LDS unit applies the lanes' updates.
//   RANK_MIX2 / RANK_MIX4  LDS-OR masks, except that every second / fourth
//                item is matched by ballots: the two methods load different
//                pipes (LDS vs VALU), and the ranking, not the writes, bounds
//                the scatter (measured on MI355X: a 256x16 scatter of 2^28
//                keys ranks and stages its tiles in 482 us with LDS-OR masks
//                alone, 395 us with every second item by ballot, against 386 us
//                for a plain copy of the same bytes).
// ---------------------------------------------------------------------------
// Tiles per range of the reduce-then-scan offset scan (k_range_totals).
// Measured on MI355X (2^28 u32): ranges of 16 tiles with 16-bit in-range
// offsets took the range and column kernels from 32 + 9 to 68 + 24 us per
// pass (four times the same-address digit-total atomics, and column loads
// over four times the ranges).
constexpr int SCAN_RANGE = 64;
constexpr int RANK_BALLOT = 0;
constexpr int RANK_LDSOR = 1;
constexpr int RANK_MIX2 = 2;
constexpr int RANK_MIX4 = 3;
// every MIX-th item of a wave is matched by ballots (0: none)
template <int RANK> constexpr int rank_mix_period() { return RANK == RANK_MIX2 ? 2 : RANK == RANK_MIX4 ? 4 : 0; }
// Default look-back rounds without progress after which a block counts the
// digits of the unpublished earlier tile itself (decoupled fallback); the
// kernel takes the value as an argument so tests can force the fallback.
constexpr int FALLBACK_ROUNDS = 24;

// 1024-thread tiles (64 KiB of staged keys) rank by ballot and keep 16-bit
// per-wave counters (a tile holds at most 16384 keys), so that two workgroups
// fit in a CU's 160 KiB of LDS; the compiler is asked for the matching
// register budget (8 waves per SIMD).
template <int BLOCK> constexpr int onesweep_waves_per_eu() { return BLOCK >= 1024 ? 8 : 1; }

template <typename K, bool HAS_VALS, int FLOATK, int BLOCK, int ITEMS, int RANK, int LOOKBACK_WINDOW>
__global__ __launch_bounds__(BLOCK, onesweep_waves_per_eu<BLOCK>()) void k_onesweep(
    K *__restrict__ keysA, K *__restrict__ keysB, uint32_t *__restrict__ valsA, uint32_t *__restrict__ valsB,
    uint32_t n, int pass, DevPlan *__restrict__ plan, uint32_t *__restrict__ lb_cur, uint32_t *__restrict__ lb_next,
    int fb_rounds, uint8_t *__restrict__ dig)
{
    static_assert(BLOCK >= RADIX && BLOCK % WAVE == 0, "one thread per digit is required");
    static_assert(BLOCK < 1024 || RANK == RANK_BALLOT, "1024-thread tiles rank by ballot");
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int SCAN_WAVES = RADIX / WAVE;
    // LOOKBACK_WINDOW == 0: the tile's global digit offsets were computed
    // beforehand (k_tile_counts + scans) and lb_cur holds them, [tile][digit].
    constexpr bool PRESCANNED = LOOKBACK_WINDOW == 0;
    // per-wave digit counters: 16-bit where ballot ranking writes them with
    // plain stores (no LDS atomics) and the tile's counts fit
    constexpr bool CNT16 = RANK == RANK_BALLOT && BLOCK >= 1024;
    static_assert(!CNT16 || TILE < 65536, "16-bit counters hold a tile's counts");
    using Cnt = typename std::conditional<CNT16, uint16_t, uint32_t>::type;

    __shared__ Cnt s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[RANK != RANK_BALLOT ? WAVES : 1][RANK != RANK_BALLOT ? RADIX : 1];
    __shared__ uint32_t s_base[RADIX];
    __shared__ uint32_t s_scan[SCAN_WAVES];
    __shared__ uint32_t s_fb[RADIX];
    __shared__ uint32_t s_fb_tile;
    __shared__ K s_keys[TILE];
    __shared__ uint32_t s_vals[HAS_VALS ? TILE : 1];

    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid >> 6;

    // Clear the next pass's look-back words for tile blockIdx.x (the grid has
    // exactly one block per tile); that region's previous pass has finished.
    if (!PRESCANNED && lb_next != nullptr && tid < RADIX) lb_next[size_t(blockIdx.x) * RADIX + tid] = 0u;

    // one packed word of per-pass flags (one dependent load before the keys);
    // pass < 0: the MSD split, whose digit position the plan holds
    const uint32_t flags = pass < 0 ? plan->msd_flags : plan->flags[pass];
    if (flags & PF_TRIVIAL) return;

    const uint32_t src = flags & PF_SRC_B;
    const K *__restrict__ kin = src ? keysB : keysA;
    K *__restrict__ kout = src ? keysA : keysB;
    const uint32_t *__restrict__ vin = src ? valsB : valsA;
    uint32_t *__restrict__ vout = src ? valsA : valsB;
    const bool xin = FLOATK && (flags & PF_XIN);
    const bool xout = FLOATK && (flags & PF_XOUT);
    const int shift = pass < 0 ? int(plan->msd_shift) : pass_shift(flags, pass);
    const uint32_t *__restrict__ gstart = pass < 0 ? plan->msd_start : plan->gstart[pass];

    // With precomputed offsets the tile order is free: give each XCD (work-
    // groups are dealt to the 8 XCDs round-robin) a contiguous range of tiles,
    // so the partial cache lines at the ends of neighbouring tiles' digit runs
    // meet in one L2 and leave it as full lines.  The look-back form keeps
    // tile = blockIdx.x: its chain wants neighbouring tiles to start together.
    const uint32_t tile = PRESCANNED ? xcd_contiguous(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t tile_base = tile * uint32_t(TILE);
    const uint32_t wave_base = tile_base + uint32_t(wave * WAVE_TILE);
    // keys of this wave that exist: items with i*64 + lane < wave_valid
    const int wave_valid = n > wave_base ? int(min(uint32_t(WAVE_TILE), n - wave_base)) : 0;

    // ---- load (issued before the LDS set-up so their latency overlaps it) ----
    // A full wave (every wave but the last tile's) loads without per-key
    // bounds tests, from one base address and immediate offsets.
    K keys[ITEMS];
    uint32_t vals[HAS_VALS ? ITEMS : 1];
    if (wave_valid == WAVE_TILE) {
        const K *__restrict__ kp = kin + wave_base + lane;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) keys[i] = kp[i * WAVE];
        // values load with the keys, so their latency hides behind the ranking
        if constexpr (HAS_VALS) {
            const uint32_t *__restrict__ vp = vin + wave_base + lane;
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) vals[i] = vp[i * WAVE];
        }
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int off = i * WAVE + lane;
            keys[i] = off < wave_valid ? kin[wave_base + off] : ~K(0);
        }
        if constexpr (HAS_VALS) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const int off = i * WAVE + lane;
                vals[i] = off < wave_valid ? vin[wave_base + off] : 0u;
            }
        }
    }
    uint32_t pre_off = 0;
    if constexpr (PRESCANNED) {
        // lb_cur: global offsets, or (lb_next != nullptr) offsets within the
        // tile's range of SCAN_RANGE tiles, lb_next then holding the ranges'
        // global bases
        if (tid < RADIX) {
            pre_off = lb_cur[size_t(tile) * RADIX + tid];
            if (lb_next != nullptr) pre_off += lb_next[size_t(tile / SCAN_RANGE) * RADIX + tid];
        }
    }

    for (int i = tid; i < WAVES * RADIX; i += BLOCK) (&s_cnt[0][0])[i] = 0u;
    if constexpr (RANK != RANK_BALLOT)
        for (int i = tid; i < WAVES * RADIX; i += BLOCK) (&s_match[0][0])[i] = 0ull;
    __syncthreads();
    if (xin) {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            if (i * WAVE + lane < wave_valid) keys[i] = key_enc<K, FLOATK>(keys[i]);
    }

    // ---- 1. per-wave digit counts ----
    // LDS-OR ranking stages keys while it ranks, so it needs the wave bases
    // first and counts with LDS atomics here.  Ballot ranking ranks first
    // with wave-local counters (their final values are the counts) and stages
    // afterwards, so it issues no counting atomics.
    Cnt *cnt = s_cnt[wave];
    // Every tile but the last is full: its loops below run without per-key
    // bounds tests (a tile-uniform branch; no barrier inside either copy).
    const bool full = n - tile_base >= uint32_t(TILE);
    // in-wave ranks (below 64 * ITEMS), two 16-bit ranks per register
    static_assert(WAVE * ITEMS < 65536, "ranks fit 16 bits");
    const bool skewed = RANK != RANK_BALLOT &&
                        wave_skewed(KeyOps<K>::digit(keys[0], shift), lane < wave_valid,
                                    KeyOps<K>::digit(keys[ITEMS / 2], shift), (ITEMS / 2) * WAVE + lane < wave_valid);
    uint32_t rank2[RANK == RANK_BALLOT ? (ITEMS + 1) / 2 : 1];
    if constexpr (RANK == RANK_BALLOT) {
        auto rank_loop = [&](auto full_c) {
            constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const uint32_t d = KeyOps<K>::digit(keys[i], shift);
                const bool valid = FULL || i * WAVE + lane < wave_valid;
                const uint64_t peers = FULL ? match_digit8(d) : match_digit8(d) & __ballot(valid);
                const uint32_t below = count_below(peers);
                const uint32_t old = cnt[d];
                if (valid && below == 0) cnt[d] = Cnt(old + uint32_t(__popcll(peers)));
                const uint32_t r = old + below;
                rank2[i / 2] = (i & 1) ? (rank2[i / 2] | (r << 16)) : r;
                __builtin_amdgcn_wave_barrier();
            }
        };
        if (full) rank_loop(std::true_type{});
        else rank_loop(std::false_type{});
        // The staging step recomputes each key's counter address instead of
        // keeping ITEMS addresses live across the block scan (an empty asm
        // that the compiler must assume rewrites the keys; no instruction).
        if constexpr (BLOCK >= 1024) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) asm volatile("" : "+v"(keys[i]));
        }
    } else {
        if (full) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) count_item(cnt, KeyOps<K>::digit(keys[i], shift), true, lane, skewed);
        } else {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                count_item(cnt, KeyOps<K>::digit(keys[i], shift), i * WAVE + lane < wave_valid, lane, skewed);
        }
    }
    __syncthreads();

    // ---- 2. per-digit wave bases, tile count, publish, scan ----
    uint32_t tile_cnt = 0, tile_start = 0;
    if (tid < RADIX) {
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint32_t c = s_cnt[w][tid];
            s_cnt[w][tid] = Cnt(tile_cnt);
            tile_cnt += c;
        }
        if constexpr (!PRESCANNED)
            lb_store(&lb_cur[size_t(tile) * RADIX + tid], (tile == 0 ? LB_INC : LB_AGG) | tile_cnt);
        // inclusive scan over the 64 digits of this wave
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
        for (int w = 0; w < WAVES; ++w) s_cnt[w][tid] = Cnt(s_cnt[w][tid] + tile_start);
    }
    __syncthreads();

    // ---- 3. rank and stage the tile in LDS in sorted order ----
    uint32_t pos[HAS_VALS ? ITEMS : 1];
    auto stage_loop = [&](auto full_c, auto peel_c) {
        constexpr bool FULL = decltype(full_c)::value;
        constexpr bool PEEL = decltype(peel_c)::value;
        if constexpr (RANK == RANK_BALLOT) {
            // cnt[d] now holds the wave's base for digit d
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                if (FULL || i * WAVE + lane < wave_valid) {
                    const uint32_t at =
                        cnt[KeyOps<K>::digit(keys[i], shift)] + ((rank2[i / 2] >> (16 * (i & 1))) & 0xFFFFu);
                    s_keys[at] = keys[i];
                    if constexpr (HAS_VALS) pos[i] = at;
                }
            }
        } else {
            uint32_t *match32 = reinterpret_cast<uint32_t *>(s_match[wave]);
            constexpr int MIX = rank_mix_period<RANK>();
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const uint32_t d = KeyOps<K>::digit(keys[i], shift);
                const bool valid = FULL || i * WAVE + lane < wave_valid;
                if (MIX && i % MIX == MIX - 1) {
                    // this item: the peers by ballots (no LDS traffic)
                    const uint64_t peers = FULL ? match_digit8(d) : match_digit8(d) & __ballot(valid);
                    const uint32_t below = count_below(peers);
                    const uint32_t old = cnt[d];
                    if (valid && below == 0) cnt[d] = Cnt(old + uint32_t(__popcll(peers)));
                    if (valid) s_keys[old + below] = keys[i];
                    if constexpr (HAS_VALS) pos[i] = old + below;
                    __builtin_amdgcn_wave_barrier();
                } else {
                    const uint32_t at = ldsor_rank_item<PEEL>(d, valid, lane, cnt, match32, s_match[wave]);
                    if (valid) s_keys[at] = keys[i];
                    if constexpr (HAS_VALS) pos[i] = at;
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        if constexpr (HAS_VALS) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                if (FULL || i * WAVE + lane < wave_valid) s_vals[pos[i]] = vals[i];
        }
    };
    if constexpr (RANK == RANK_BALLOT) {
        if (full) stage_loop(std::true_type{}, std::false_type{});
        else stage_loop(std::false_type{}, std::false_type{});
    } else if (full) {
        if (skewed) stage_loop(std::true_type{}, std::true_type{});
        else stage_loop(std::true_type{}, std::false_type{});
    } else {
        if (skewed) stage_loop(std::false_type{}, std::true_type{});
        else stage_loop(std::false_type{}, std::false_type{});
    }

    // ---- 4. decoupled look-back over a window of earlier tiles ----
    // A round reads LOOKBACK_WINDOW words at once; after meeting an
    // unpublished word the thread polls only that word (with a pause) until it
    // is published, so waiting costs one load per round, not a window.
    if constexpr (PRESCANNED) {
        if (tid < RADIX) s_base[tid] = pre_off - tile_start;
    } else {
        uint32_t excl = 0;
        int64_t p = int64_t(tile) - 1; // newest earlier tile not yet added
        bool done = tid >= RADIX || tile == 0;
        bool polling = false;
        int idle = 0;                  // rounds in a row without progress
        while (true) {
            if (!done) {
                if (polling) {
                    const uint32_t w = lb_load(&lb_cur[size_t(p) * RADIX + tid]);
                    const uint32_t st = w & ~LB_VALUE;
                    if (st != 0u) {
                        excl += w & LB_VALUE;
                        done = st == LB_INC;
                        p -= 1;
                        polling = false;
                        idle = 0;
                    } else {
                        ++idle;
                    }
                } else {
                    const int avail = p + 1 < LOOKBACK_WINDOW ? int(p + 1) : LOOKBACK_WINDOW;
                    uint32_t v[LOOKBACK_WINDOW > 0 ? LOOKBACK_WINDOW : 1];
#pragma unroll
                    for (int k = 0; k < LOOKBACK_WINDOW; ++k)
                        v[k] = k < avail ? lb_load(&lb_cur[size_t(p - k) * RADIX + tid]) : 0u;
                    int used = 0;
#pragma unroll
                    for (int k = 0; k < LOOKBACK_WINDOW; ++k) {
                        if (!done && used == k && k < avail) {
                            const uint32_t st = v[k] & ~LB_VALUE;
                            if (st != 0u) {
                                excl += v[k] & LB_VALUE;
                                used = k + 1;
                                done = st == LB_INC;
                            }
                        }
                    }
                    p -= used;
                    polling = !done && used < avail; // stopped at an unpublished word
                    idle = used ? 0 : idle + 1;
                }
            }
            if (!__syncthreads_or(!done)) break;
            if (__syncthreads_or(!done && idle >= fb_rounds)) {
                // count the newest unpublished tile's digits here
                if (tid == 0) s_fb_tile = 0u;
                for (int i = tid; i < RADIX; i += BLOCK) s_fb[i] = 0u;
                __syncthreads();
                if (!done && idle >= fb_rounds) atomicMax(&s_fb_tile, uint32_t(p));
                __syncthreads();
                const uint32_t q = s_fb_tile;
                const uint32_t qbase = q * uint32_t(TILE);
                const uint32_t qn = min(uint32_t(TILE), n - qbase);
                for (uint32_t j = tid; j < qn; j += BLOCK) {
                    const K k = kin[qbase + j];
                    atomicAdd(&s_fb[KeyOps<K>::digit(xin ? key_enc<K, FLOATK>(k) : k, shift)], 1u);
                }
                __syncthreads();
                if (!done && uint32_t(p) == q) {
                    excl += s_fb[tid];
                    p -= 1;
                    idle = 0;
                    polling = false;
                    done = p < 0; // counted every earlier tile
                }
                __syncthreads();
            } else if (!done && polling) {
                __builtin_amdgcn_s_sleep(8);
            }
        }
        if (tid < RADIX) {
            if (tile > 0) lb_store(&lb_cur[size_t(tile) * RADIX + tid], LB_INC | (excl + tile_cnt));
            s_base[tid] = gstart[tid] + excl - tile_start;
        }
    }
    __syncthreads();

    // ---- 5. write the tile out, each digit's run contiguous ----
    // With `dig` (reduce-then-scan form) each key's digit of the next pass
    // that moves keys is also written, one byte at the key's new position, so
    // that pass counts its tiles from n bytes instead of re-reading the keys.
    const uint32_t next_pass = (flags >> PF_NEXT_SHIFT) & PF_NEXT_NONE;
    const bool wdig = PRESCANNED && dig != nullptr && next_pass != PF_NEXT_NONE;
    const int next_shift = int(next_pass) * RADIX_BITS;
    const int tile_valid = n > tile_base ? int(min(uint32_t(TILE), n - tile_base)) : 0;
    auto write_loop = [&](auto full_c) {
        constexpr bool FULL = decltype(full_c)::value;
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int j = i * BLOCK + tid;
            if (FULL || j < tile_valid) {
                const K k = s_keys[j];
                const uint32_t dst = s_base[KeyOps<K>::digit(k, shift)] + uint32_t(j);
                kout[dst] = xout ? key_dec<K, FLOATK>(k) : k;
                if constexpr (HAS_VALS) vout[dst] = s_vals[j];
                if (wdig) dig[dst] = uint8_t(KeyOps<K>::digit(k, next_shift));
            }
        }
    };
    if (full) write_loop(std::true_type{});
    else write_loop(std::false_type{});
}

// ---------------------------------------------------------------------------
// Reduce-then-scan form of a pass (no waits between workgroups):
//   k_tile_counts    per-tile digit counts, [tile][digit]
//   k_range_totals   per range of SCAN_RANGE tiles, the sum of its counts,
//                    and each tile's offset within the range over its counts
//   k_scan_columns   exclusive scan of the range sums over ranges, seeded with
//                    the digit's global start: the ranges' global bases
// followed by k_onesweep<..., LOOKBACK_WINDOW = 0> adding its range's base to
// its in-range offsets.
// It reads the keys twice per pass (12 B per u32 key instead of 8) and has
// no chain of dependent loads between tiles.
// ---------------------------------------------------------------------------

// Keys of one tile for counting (order does not matter): with `vec` (the
// buffer is 16-byte aligned) a full tile is read in 16-byte vectors, thread t
// taking vectors t, t + BLOCK, ...; key slot i*BLOCK + t of the scalar form
// is replaced by a permutation of the same keys.  Missing keys read as 0 and
// are masked by the caller (slot index >= cnt).
template <typename K, int BLOCK, int ITEMS>
__device__ inline void load_tile_keys(const K *__restrict__ p, uint32_t cnt, bool vec, K (&k)[ITEMS])
{
    constexpr int VK = 16 / int(sizeof(K));
    constexpr int NV = ITEMS / VK;
    using V = typename Vec16<K>::type;
    const int tid = threadIdx.x;
    if (vec && cnt == uint32_t(BLOCK * ITEMS)) {
        const V *pv = reinterpret_cast<const V *>(p);
        V v[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = pv[i * BLOCK + tid];
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
            for (int e = 0; e < VK; ++e) k[i * VK + e] = v[i][e];
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t j = uint32_t(i * BLOCK + tid);
            k[i] = j < cnt ? p[j] : K(0);
        }
    }
}

// Count one digit per LDS atomic; the lanes that share the first lane's
// digit, when many do, add once (count_peeled).
__device__ inline void count_digit(uint32_t *sh, uint32_t d, bool valid, int lane)
{
    count_peeled(sh, d, valid, lane);
}

// Per-tile digit counts of a pass (16-bit: a tile holds fewer than 65536
// keys).  When the previous pass that moved keys
// wrote this pass's digits as bytes (`dig`, every pass after the first that
// moves keys), the tile's TILE bytes are read (thread t: bytes t*ITEMS ..,
// in 4-byte words); otherwise the keys are.
template <typename K, int FLOATK, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_tile_counts(const K *__restrict__ keysA, const K *__restrict__ keysB,
                                                       uint32_t n, int pass, const DevPlan *__restrict__ plan,
                                                       uint16_t *__restrict__ counts, int vec,
                                                       const uint8_t *__restrict__ dig)
{
    static_assert(ITEMS % 4 == 0, "digit bytes are read in 4-byte words");
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(TILE < 65536, "tile counts are 16-bit");
    __shared__ uint32_t sh[RADIX];
    const uint32_t flags = plan->flags[pass];
    if (flags & (PF_TRIVIAL | PF_PRECOUNTED | PF_PRE0)) return;
    const K *__restrict__ kin = (flags & PF_SRC_B) ? keysB : keysA;
    const bool xin = FLOATK && (flags & PF_XIN);
    const int shift = pass_shift(flags, pass);
    const int tid = threadIdx.x, lane = tid & (WAVE - 1);
    const uint32_t base = blockIdx.x * uint32_t(TILE);
    if (base >= n) return; // a grid beyond the tiles (every launch sizes it to them): nothing to count
    const uint32_t cnt = min(uint32_t(TILE), n - base);
    if (dig == nullptr || (flags & PF_FIRST)) {
        // keys: the loads go out first, so their latency covers the LDS set-up
        K k[ITEMS];
        load_tile_keys<K, BLOCK, ITEMS>(kin + base, cnt, vec, k);
        for (int i = tid; i < RADIX; i += BLOCK) sh[i] = 0u;
        __syncthreads();
        if (cnt == uint32_t(TILE)) { // a full tile: no per-key bounds test
#pragma unroll
            for (int i = 0; i < ITEMS; ++i)
                count_digit(sh, KeyOps<K>::digit(xin ? key_enc<K, FLOATK>(k[i]) : k[i], shift), true, lane);
        } else {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const uint32_t j = uint32_t(i * BLOCK + tid);
                count_digit(sh, KeyOps<K>::digit(xin ? key_enc<K, FLOATK>(k[i]) : k[i], shift), j < cnt, lane);
            }
        }
        __syncthreads();
        for (int i = tid; i < RADIX; i += BLOCK) counts[size_t(blockIdx.x) * RADIX + i] = uint16_t(sh[i]);
        return;
    }
    for (int i = tid; i < RADIX; i += BLOCK) sh[i] = 0u;
    __syncthreads();
    {
        const uint32_t first = uint32_t(tid * ITEMS); // this thread's bytes [first, first + ITEMS)
        uint32_t w[ITEMS / 4];
        if (cnt == uint32_t(TILE)) {
            const uint32_t *pw = reinterpret_cast<const uint32_t *>(dig + base + first);
#pragma unroll
            for (int i = 0; i < ITEMS / 4; ++i) w[i] = pw[i];
        } else {
#pragma unroll
            for (int i = 0; i < ITEMS / 4; ++i) {
                uint32_t x = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const uint32_t j = first + uint32_t(4 * i + b);
                    if (j < cnt) x |= uint32_t(dig[base + j]) << (8 * b);
                }
                w[i] = x;
            }
        }
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            count_digit(sh, (w[i / 4] >> (8 * (i % 4))) & 0xFFu, first + uint32_t(i) < cnt, lane);
        __syncthreads();
        for (int i = tid; i < RADIX; i += BLOCK) counts[size_t(blockIdx.x) * RADIX + i] = uint16_t(sh[i]);
    }
}

// Spacing of per-digit accumulators that take atomics from many workgroups:
// one 128-byte line each, so the atomics do not queue behind each other.
constexpr int LINE_WORDS = 32;
// Slots for OR(keys) / OR(~keys) of the first read (k_upsweep): atomics
// spread over ORSLOTS lines.
constexpr int ORSLOTS = 64;
template <typename K> constexpr int OR_STRIDE = 128 / int(sizeof(K)); // one 128-byte line per slot

// totals[r][d] = sum of counts[t][d] over the SCAN_RANGE tiles of range r,
// offsets[t][d] = the sum of the range's earlier counts (the tile's offset
// within the range), and the pass's digit totals dtot[d * LINE_WORDS] +=
// totals[r][d].  A PF_PRECOUNTED pass reads counts_pre (the upsweep's).
static __global__ __launch_bounds__(RADIX) void k_range_totals(const uint16_t *__restrict__ counts_own,
                                                               const uint16_t *__restrict__ counts_pre,
                                                               uint32_t *__restrict__ offsets, uint32_t tiles,
                                                               uint32_t *__restrict__ totals,
                                                               uint32_t *__restrict__ dtot, int pass,
                                                               const DevPlan *__restrict__ plan)
{
    const uint32_t flags = plan->flags[pass];
    if (flags & PF_TRIVIAL) return;
    const uint16_t *__restrict__ counts = (flags & PF_PRECOUNTED) ? counts_pre : counts_own;
    const uint32_t r = blockIdx.x, d = threadIdx.x;
    const uint32_t t0 = r * SCAN_RANGE, t1 = min(tiles, t0 + SCAN_RANGE);
    uint32_t c[SCAN_RANGE];
#pragma unroll
    for (int k = 0; k < SCAN_RANGE; ++k) c[k] = t0 + k < t1 ? uint32_t(counts[size_t(t0 + k) * RADIX + d]) : 0u;
    // each tile's offset within the range
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_RANGE; ++k) {
        if (t0 + k < t1) offsets[size_t(t0 + k) * RADIX + d] = sum;
        sum += c[k];
    }
    totals[size_t(r) * RADIX + d] = sum;
    if (sum) atomicAdd(&dtot[d * LINE_WORDS], sum);
}

// One workgroup per digit d: the digit's global start (ComputeStartOfBins,
// RadixSortLsd.cs:112-125: the sum of the smaller digits' totals), then the
// exclusive scan of column d of the range totals, seeded with that start.
static __global__ __launch_bounds__(1024) void k_scan_columns(uint32_t *__restrict__ totals, uint32_t nranges,
                                                       const uint32_t *__restrict__ dtot, int pass,
                                                       DevPlan *__restrict__ plan)
{
    __shared__ uint32_t s_w[1024 / WAVE];
    if (plan->flags[pass] & PF_TRIVIAL) return;
    const uint32_t d = blockIdx.x, tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    // global start of digit d
    uint32_t v = (tid < d) ? dtot[tid * LINE_WORDS] : 0u;
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) s_w[wave] = v;
    __syncthreads();
    uint32_t start = 0;
    for (int w = 0; w < 1024 / WAVE; ++w) start += s_w[w];
    if (tid == 0) plan->gstart[pass][d] = start;
    __syncthreads();
    // column scan: thread tid owns ranges [tid*per, (tid+1)*per)
    const uint32_t per = (nranges + 1023) / 1024;
    const uint32_t r0 = min(nranges, tid * per), r1 = min(nranges, r0 + per);
    uint32_t local = 0;
    for (uint32_t r = r0; r < r1; ++r) local += totals[size_t(r) * RADIX + d];
    uint32_t x = local;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= uint32_t(off)) x += y;
    }
    if (lane == WAVE - 1) s_w[wave] = x;
    __syncthreads();
    uint32_t run = start + x - local;
    for (uint32_t w = 0; w < wave; ++w) run += s_w[w];
    for (uint32_t r = r0; r < r1; ++r) {
        const size_t at = size_t(r) * RADIX + d;
        const uint32_t c = totals[at];
        totals[at] = run;
        run += c;
    }
}

// The three scan kernels in one, for sorts of few tiles (the plan's digit
// starts come from a full histogram, k_histogram + k_plan): workgroup d
// scans digit d's column of the tile counts, the tiles split over its
// threads, and writes the tiles' 32-bit global offsets to `offsets`.
// Measured on MI355X: faster than the five-launch pass up to 512 tiles (u32
// 2^20: 0.109 -> 0.092 ms, u64 2^20: 0.240 -> 0.199 ms), slower from 1024
// (the all-digit histogram and the column loads outgrow the launches saved).
constexpr uint32_t SCAN_DIGIT_MAX_TILES = 512;
static __global__ __launch_bounds__(1024) void k_scan_digit(const uint16_t *__restrict__ counts,
                                                     uint32_t *__restrict__ offsets, uint32_t tiles, int pass,
                                                     const DevPlan *__restrict__ plan)
{
    __shared__ uint32_t s_w[1024 / WAVE];
    if (plan->flags[pass] & PF_TRIVIAL) return;
    const uint32_t d = blockIdx.x, tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid / WAVE;
    const uint32_t per = (tiles + 1023) / 1024;
    const uint32_t t0 = min(tiles, tid * per), t1 = min(tiles, t0 + per);
    uint32_t local = 0;
    for (uint32_t t = t0; t < t1; ++t) local += counts[size_t(t) * RADIX + d];
    uint32_t x = local;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= uint32_t(off)) x += y;
    }
    if (lane == WAVE - 1) s_w[wave] = x;
    __syncthreads();
    uint32_t run = plan->gstart[pass][d] + x - local;
    for (uint32_t w = 0; w < wave; ++w) run += s_w[w];
    for (uint32_t t = t0; t < t1; ++t) {
        const size_t at = size_t(t) * RADIX + d;
        const uint32_t c = counts[at];
        offsets[at] = run;
        run += c;
    }
}

// ---------------------------------------------------------------------------
// k_upsweep: the first read of a reduce-then-scan sort.  Per tile it writes
// the digit-0 counts (pass 0's reduce step, [tile][digit]) and folds the keys
// into OR(keys) and OR(~keys): a bit varies across the keys iff it is set in
// both, so digit position p is trivial (one value for every key: a stable
// identity pass) iff no bit of its byte varies.  The two ORs go to one of
// ORSLOTS word pairs (by block) to keep atomics on one address few.
// ---------------------------------------------------------------------------

template <typename K, int FLOATK, int BLOCK, int ITEMS, bool TOP = false>
__global__ __launch_bounds__(BLOCK) void k_upsweep(const K *__restrict__ keys, uint32_t n,
                                                   uint16_t *__restrict__ counts, K *__restrict__ orslots, int vec,
                                                   uint16_t *__restrict__ counts_top = nullptr)
{
    constexpr int TILE = BLOCK * ITEMS;
    static_assert(TILE < 65536, "tile counts are 16-bit");
    static_assert(!TOP || (sizeof(K) == 4 && FLOATK == 0), "top-byte counts serve the MSD-first u32 form");
    constexpr int WAVES = BLOCK / WAVE;
    __shared__ uint32_t sh[RADIX];
    __shared__ uint32_t sh_top[TOP ? RADIX : 1];
    __shared__ K s_or[WAVES], s_nor[WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    const uint32_t base = blockIdx.x * uint32_t(TILE);
    if (base >= n) return; // a grid beyond the tiles (every launch sizes it to them): nothing to count
    const uint32_t cnt = min(uint32_t(TILE), n - base);
    // the loads go out first, so their latency covers the LDS set-up
    K k[ITEMS];
    load_tile_keys<K, BLOCK, ITEMS>(keys + base, cnt, vec, k);
    for (int i = tid; i < RADIX; i += BLOCK) {
        sh[i] = 0u;
        if constexpr (TOP) sh_top[i] = 0u;
    }
    __syncthreads();
    K vor = 0, vnor = 0;
    auto count_loop = [&](auto full_c) {
        constexpr bool FULL = decltype(full_c)::value; // a full tile: no per-key bounds test
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t j = uint32_t(i * BLOCK + tid);
            const bool valid = FULL || j < cnt;
            const K key = FLOATK ? key_enc<K, FLOATK>(k[i]) : k[i];
            if (valid) {
                vor |= key;
                vnor |= ~key;
            }
            count_digit(sh, KeyOps<K>::digit(key, 0), valid, lane);
            if constexpr (TOP) count_digit(sh_top, KeyOps<K>::digit(key, 24), valid, lane);
        }
    };
    if (cnt == uint32_t(TILE)) count_loop(std::true_type{});
    else count_loop(std::false_type{});
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
    for (int i = tid; i < RADIX; i += BLOCK) {
        counts[size_t(blockIdx.x) * RADIX + i] = uint16_t(sh[i]);
        if constexpr (TOP) counts_top[size_t(blockIdx.x) * RADIX + i] = uint16_t(sh_top[i]);
    }
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

// k_plan_rts: trivial passes from the varying bits; buffer selection and the
// float transform's first/last pass.  Digit starts come later, per pass, from
// that pass's counts (k_scan_ranges).  With `hybrid` (4-byte unsigned keys
// without values whose upsweep also counted the top byte; the launch
// sequence of hpcs_hybrid.hpp, whose LSD loop has pass slot 3 only): if a bit
// from 23 up varies, the MSD-first form -- slot 3 splits A to B by the 8 bits
// that end at the top varying bit (at most the top byte, at least bits
// 16..23; counted by the upsweep when that is the top byte, else by the
// loop's counting kernel), the bucket stages bring the keys back to A sorted;
// otherwise the LSD passes over the (at most three) varying bytes, lowest
// first: slot 3 (A to B; the upsweep's byte-0 counts when that is byte 0),
// the segmented pass over one segment of all keys (B to A), the first
// overflow pass (A to B), and a copy back after an odd number of them.
template <typename K>
__global__ void k_plan_rts(const K *__restrict__ orslots, DevPlan *__restrict__ plan, uint64_t n, int floatk,
                           int hybrid = 0)
{
    if (threadIdx.x != 0) return;
    constexpr int P = int(sizeof(K));
    K a = 0, b = 0;
    for (int s = 0; s < ORSLOTS; ++s) {
        a |= orslots[s * OR_STRIDE<K>];
        b |= orslots[s * OR_STRIDE<K> + 1];
    }
    const K varying = a & b;
    if (P == 4 && hybrid && !floatk) {
        uint32_t mode = HYB_NONE, s1 = 0, s2 = HYB_SKIP, ov0 = HYB_SKIP, ov1 = HYB_SKIP, copy = 0, pre = 0;
        if (n > 1 && varying != 0) {
            const int top = 63 - __builtin_clzll(uint64_t(varying));
            if (top >= 23) {
                mode = HYB_MSD;
                s1 = uint32_t(min(24, top - 7));
                s2 = s1 - 8;
                ov0 = 0;
                ov1 = 8;
                pre = s1 == 24 ? PF_PRECOUNTED : 0u;
            } else {
                mode = HYB_LSD;
                uint32_t v[3], k = 0;
                for (uint32_t b = 0; b < 3; ++b)
                    if ((varying >> (8 * b)) & K(0xFF)) v[k++] = 8 * b;
                s1 = v[0];
                pre = s1 == 0 ? PF_PRE0 : 0u;
                if (k >= 2) s2 = v[1];
                if (k >= 3) ov0 = v[2];
                copy = k & 1u;
            }
        }
        for (int p = 0; p < P; ++p) {
            plan->trivial[p] = (p == 3 && mode != HYB_NONE) ? 0u : 1u;
            plan->src[p] = 0u;
            plan->xin[p] = 0u;
            plan->xout[p] = 0u;
        }
        plan->copyback = copy;
        plan->hybrid = mode;
        plan->hyb_shift1 = s1;
        plan->hyb_shift2 = s2;
        plan->hyb_ov_shift[0] = ov0;
        plan->hyb_ov_shift[1] = ov1;
        pack_plan_flags(plan, P);
        for (int p = 0; p < P; ++p) plan->flags[p] |= PF_NEXT_NONE << PF_NEXT_SHIFT;
        if (mode != HYB_NONE) plan->flags[3] |= PF_FIRST | PF_SHIFT | (s1 << PF_SHIFT_POS) | pre;
        return;
    }
    plan->hybrid = 0u;
    uint32_t cur = 0;
    int first = -1, last = -1;
    for (int p = 0; p < P; ++p) {
        const uint32_t t = (n <= 1 || ((varying >> (8 * p)) & K(0xFF)) == 0) ? 1u : 0u;
        plan->trivial[p] = t;
        plan->src[p] = cur;
        if (!t) {
            cur ^= 1u;
            if (first < 0) first = p;
            last = p;
        }
    }
    for (int p = 0; p < P; ++p) {
        plan->xin[p] = (floatk && p == first) ? 1u : 0u;
        plan->xout[p] = (floatk && p == last) ? 1u : 0u;
    }
    plan->copyback = cur;
    pack_plan_flags(plan, P);
    uint32_t next = PF_NEXT_NONE;
    for (int p = P - 1; p >= 0; --p) {
        plan->flags[p] |= next << PF_NEXT_SHIFT;
        if (!plan->trivial[p]) next = uint32_t(p);
    }
    if (first >= 0) plan->flags[first] |= PF_FIRST;
}

// ---------------------------------------------------------------------------
// k_copyback: if the result sits in buffer B, copy it to A.
// ---------------------------------------------------------------------------
template <typename K>
__global__ void k_copyback(K *__restrict__ keysA, const K *__restrict__ keysB, uint32_t *__restrict__ valsA,
                           const uint32_t *__restrict__ valsB, uint64_t n, const DevPlan *__restrict__ plan)
{
    if (!plan->copyback) return;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        keysA[i] = keysB[i];
        if (valsA) valsA[i] = valsB[i];
    }
}

// ---------------------------------------------------------------------------
// N-bit digit histograms (Histogram11bitComponents, HistogramNbitComponents):
// digit d of a key is (key >> (d * bits)) & (2^bits - 1).  The tables of all
// digits sit in dynamic LDS when they fit (`in_lds`), else the counts go
// straight to the global result.  Grid-stride; a wave whose valid lanes share
// a digit adds once.
// ---------------------------------------------------------------------------
template <typename K>
__global__ __launch_bounds__(256) void k_histogram_nbit(const K *__restrict__ keys, uint64_t n, int bits, int digits,
                                                        int in_lds, uint32_t *__restrict__ counts)
{
    extern __shared__ uint32_t sh_nbit[];
    const uint32_t bins = 1u << bits, mask = bins - 1u;
    const uint32_t table = uint32_t(digits) * bins;
    uint32_t *tab = in_lds ? sh_nbit : counts;
    if (in_lds) {
        for (uint32_t i = threadIdx.x; i < table; i += blockDim.x) sh_nbit[i] = 0u;
        __syncthreads();
    }
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const uint64_t start = uint64_t(blockIdx.x) * blockDim.x;
    const int lane = threadIdx.x & (WAVE - 1);
    for (uint64_t base = start; base < n; base += stride) {
        const uint64_t i = base + threadIdx.x;
        const bool valid = i < n;
        const K k = valid ? keys[i] : K(0);
        const uint64_t vm = __ballot(valid);
        const int first = __builtin_ctzll(vm);
        for (int d = 0; d < digits; ++d) {
            const uint32_t v = uint32_t(k >> (d * bits)) & mask;
            const uint32_t v0 = __shfl(v, first);
            if (__ballot(valid && v == v0) == vm) {
                if (lane == first) atomicAdd(&tab[uint32_t(d) * bins + v0], uint32_t(__popcll(vm)));
            } else if (valid) {
                atomicAdd(&tab[uint32_t(d) * bins + v], 1u);
            }
        }
    }
    if (in_lds) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < table; i += blockDim.x)
            if (sh_nbit[i]) atomicAdd(&counts[i], sh_nbit[i]);
    }
}

// ---------------------------------------------------------------------------
// Records: extract (radix key, index) pairs, and gather records by index.
// ---------------------------------------------------------------------------
// With copy16 (records 16-byte aligned, size a multiple of 16) each thread
// also copies its record there in 16-byte vectors: the extraction reads the
// records' lines anyway.
template <typename K>
__global__ void k_extract_keys(const unsigned char *__restrict__ recs, uint64_t n, int rec_size, int key_offset,
                               int kind, K *__restrict__ keys, uint32_t *__restrict__ idx, u32x4 *__restrict__ copy16)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const unsigned char *p = recs + i * uint64_t(rec_size) + key_offset;
        if (copy16) {
            const uint32_t v = uint32_t(rec_size) >> 4;
            const u32x4 *src = reinterpret_cast<const u32x4 *>(recs) + i * v;
            for (uint32_t j = 0; j < v; ++j) copy16[i * v + j] = src[j];
        }
        K k;
        if ((reinterpret_cast<uintptr_t>(p) & (sizeof(K) - 1)) == 0) {
            k = *reinterpret_cast<const K *>(p);
        } else {
            k = 0;
            for (int b = 0; b < int(sizeof(K)); ++b) k |= K(p[b]) << (8 * b);
        }
        keys[i] = kind == KIND_FLOAT ? KeyOps<K>::to_key(k) : kind == KIND_INT ? K(k ^ KeyOps<K>::SIGN) : k;
        if (idx) idx[i] = uint32_t(i);
    }
}

// One thread per 4-byte word of output when rec_size and the buffer allow it
// (words4), one per byte otherwise.
static __global__ void k_gather_records(const unsigned char *__restrict__ in, unsigned char *__restrict__ out,
                                 const uint32_t *__restrict__ idx, uint64_t n, int rec_size, int words4)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    if (words4 == 2) {
        // 16-byte aligned records of 16 * v bytes: one 16-byte vector per thread
        const uint32_t v = uint32_t(rec_size) >> 4;
        const uint64_t total = n * v;
        const u32x4 *in4 = reinterpret_cast<const u32x4 *>(in);
        u32x4 *out4 = reinterpret_cast<u32x4 *>(out);
        for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
            const uint64_t r = v == 1u ? t : t / v;
            out4[t] = in4[uint64_t(idx[r]) * v + (t - r * v)];
        }
    } else if (words4) {
        const int words = rec_size >> 2;
        const uint64_t total = n * uint64_t(words);
        for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
            const uint64_t r = t / words;
            const int w = int(t - r * words);
            reinterpret_cast<uint32_t *>(out)[t] =
                reinterpret_cast<const uint32_t *>(in + uint64_t(idx[r]) * rec_size)[w];
        }
    } else {
        const uint64_t total = n * uint64_t(rec_size);
        for (uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
            const uint64_t r = t / rec_size;
            const int b = int(t - r * rec_size);
            out[t] = in[uint64_t(idx[r]) * rec_size + b];
        }
    }
}

static __global__ void k_iota_u32(uint32_t *__restrict__ out, uint64_t n)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = uint32_t(i);
}

} // namespace hpcs
