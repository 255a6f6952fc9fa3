// hpcs_topk.hpp -- radix top-k: the m largest keys of an array, in a fixed
// sequence of launches with every decision taken on the device.
//
// The reference selects in place with 16-bit MSD digits, gathering the bin
// that holds index k and recursing into it (SelectionTopKRadix.cs:96-133).
// Here the same MSD descent runs with 11-bit digits over a device-resident
// state instead of in-place moves:
//   k_topk_hist     level-0 digit counts of the whole input (one read)
//   per level l:
//     k_topk_select   one workgroup: the bin b holding the m-th largest key,
//                     how many of it are still needed (k_rem), the key prefix
//     k_topk_compact  keys above b are certain: staged in LDS and appended to
//                     the output; keys in b are the next level's candidates
//                     (their next digit is counted on the way); keys below b
//                     go to the optional `rest` output.  If every key of b is
//                     needed, b is taken whole and the descent stops.
//   k_topk_finish   after the last level the m-th largest key T is known:
//                   k_rem copies of T close the output, the surplus copies of
//                   T go to `rest`
// Large inputs without a `rest` output first take a sample and filter:
//   k_topk_sample   level-0 digit counts of an evenly spaced sample (2^16
//                   keys)
//   k_topk_filter   one read: from the sample, a level-0 bin b_est below the
//                   sample's (2 m S / n + 16)-th largest key; keys of bins >=
//                   b_est are kept as the candidate list and counted (the
//                   level-0 counts of those bins)
// and level 0 then descends over the kept keys alone.  If fewer than m were
// kept (a sample that misrepresents the input), k_topk_hist counts the keys
// below b_est and level 0 reads the whole input, as without the filter.
// Output positions come from atomics, so the selected keys are then sorted
// (one workgroup: by counting ranks up to TOPK_RANK_SORT keys, by an LSD
// radix sort in LDS up to TOPK_BLOCK_SORT keys; beyond, the full radix sort)
// and written back as values, descending or ascending.
#pragma once
#include "hpcs_kernels.hpp"
#include "hpcs_segments.hpp"

namespace hpcs {

constexpr int TOPK_BITS = 11;
constexpr int TOPK_BINS = 1 << TOPK_BITS;
constexpr int TOPK_BLOCK = 256;
constexpr int TOPK_HIST_BLOCK = 1024;
constexpr int TOPK_STAGE = 512; // keys per wave and list staged in LDS by the compaction
constexpr int TOPK_MAX_LEVELS = 6;
constexpr int TOPK_SORT_BLOCK = 1024;
constexpr int TOPK_SORT_ITEMS = 16;
constexpr int TOPK_BLOCK_SORT = TOPK_SORT_BLOCK * TOPK_SORT_ITEMS; // largest selection sorted in one workgroup
constexpr int TOPK_RANK_SORT = 256; // up to this many keys, ranked by counting instead

// Keys per thread and tile: 64 bytes per thread (four 16-byte vectors).
template <typename K> struct TopkTile {
    static constexpr int ITEMS = 64 / int(sizeof(K));
    static constexpr int KEYS = TOPK_BLOCK * ITEMS;
};

// Level l covers key bits [shift(l), shift(l) + width(l)), from the top; the
// last level takes what is left (u32: 11+11+10, u64: 5x11+9).
template <typename K> struct TopkLevels {
    static constexpr int BITS = 8 * int(sizeof(K));
    static constexpr int LEVELS = (BITS + TOPK_BITS - 1) / TOPK_BITS;
    __host__ __device__ static constexpr int shift(int l)
    {
        return BITS - TOPK_BITS * l > TOPK_BITS ? BITS - TOPK_BITS * (l + 1) : 0;
    }
    __host__ __device__ static constexpr int width(int l)
    {
        return BITS - TOPK_BITS * l > TOPK_BITS ? TOPK_BITS : BITS - TOPK_BITS * l;
    }
    __device__ static inline uint32_t digit(K u, int l)
    {
        return uint32_t(u >> shift(l)) & ((1u << width(l)) - 1u);
    }
};
static_assert(TopkLevels<uint64_t>::LEVELS <= TOPK_MAX_LEVELS, "level table");

// Selection state of level l: the keys still needed from the candidates
// (k_rem), and after the level's threshold bin is chosen, the bin, its count
// and the key prefix fixed so far.
struct TopkSel {
    uint32_t k_rem;
    uint32_t bin;
    uint32_t bin_count;
    uint32_t pad;
    uint64_t prefix;
    uint64_t pad2;
};

// Level l reads counts hist[l % 3] and candidate count n_cand[l % 3], adds
// into the (l + 1) % 3 entries and clears the (l + 2) % 3 entries, which the
// previous level read and the next level fills: no launch is needed between
// levels to reset them.
struct TopkState {
    uint32_t hist[3][TOPK_BINS];
    uint32_t n_cand[4];
    uint32_t n_out;      // keys appended to the output
    uint32_t n_rest;     // keys appended to rest
    uint32_t done;       // 1 + the level whose bin was taken whole; 0 while descending
    uint32_t pad[9];
    TopkSel sel[TOPK_MAX_LEVELS + 1]; // sel[l + 1] written by level l
    // sampled filter (inputs of at least TOPK_FILTER_MIN keys, no rest)
    uint32_t shist[TOPK_BINS]; // level-0 counts of the sample
    uint32_t n_filt;           // keys kept by the filter (level-0 bin >= b_est)
    uint32_t b_est;
    uint32_t pad3[2];
};
constexpr uint32_t TOPK_FILTER_MIN = 1u << 20; // smallest input that is sampled
constexpr int TOPK_SAMPLE_BLOCKS = 64;
constexpr int TOPK_SAMPLE_ITEMS = 4;
constexpr uint32_t TOPK_SAMPLE = uint32_t(TOPK_SAMPLE_BLOCKS) * TOPK_BLOCK * TOPK_SAMPLE_ITEMS; // 2^16 keys

// Order key of a value: the key-kind transform (float: sign flip / all-bits
// flip; signed: sign flip; unsigned: the value itself).
template <typename K, int FLOATK> __device__ inline K order_key(K raw)
{
    return key_enc<K, FLOATK>(raw);
}
template <typename K, int FLOATK> __device__ inline K value_of(K u)
{
    return key_dec<K, FLOATK>(u);
}
// Sort key of an order key: ascending sort keys give the requested order.
template <typename K> __device__ inline K sort_key(K u, int desc)
{
    return desc ? K(~u) : u;
}

// Add one to s_hist[d] for each valid lane; a wave whose valid lanes share a
// digit adds once (sorted and constant inputs).
__device__ inline void wave_count(uint32_t *s_hist, uint32_t d, bool valid)
{
    const uint64_t vm = __ballot(valid);
    if (vm == 0) return;
    const int first = __builtin_ctzll(vm);
    const uint32_t d0 = __shfl(d, first);
    const uint64_t same = __ballot(valid && d == d0);
    if (same == vm) {
        if ((threadIdx.x & (WAVE - 1)) == uint32_t(first)) atomicAdd(&s_hist[d0], uint32_t(__popcll(vm)));
    } else if (valid) {
        atomicAdd(&s_hist[d], 1u);
    }
}

// Level-0 counts over the whole input (grid-stride over tiles).
// With `sampled` (the filter ran before): nothing to do when the filter kept
// at least m keys; otherwise only the keys below the filter's bin are
// counted, the filter having counted the others.
template <typename K, int FLOATK>
__global__ __launch_bounds__(TOPK_HIST_BLOCK) void k_topk_hist(const K *__restrict__ in, uint32_t n, int vec,
                                                               TopkState *__restrict__ st, int sampled, uint32_t m)
{
    using LV = TopkLevels<K>;
    constexpr int ITEMS = TopkTile<K>::ITEMS, TILE = TOPK_HIST_BLOCK * ITEMS;
    __shared__ uint32_t s_hist[TOPK_BINS];
    if (sampled && st->n_filt >= m) return;
    const uint32_t below = sampled ? st->b_est : uint32_t(TOPK_BINS);
    for (int i = threadIdx.x; i < TOPK_BINS; i += TOPK_HIST_BLOCK) s_hist[i] = 0u;
    __syncthreads();
    for (uint32_t base = blockIdx.x * uint32_t(TILE); base < n; base += gridDim.x * uint32_t(TILE)) {
        const uint32_t cnt = min(uint32_t(TILE), n - base);
        const bool full = vec && cnt == uint32_t(TILE);
        K u[ITEMS];
        load_tile_keys<K, TOPK_HIST_BLOCK, ITEMS>(in + base, cnt, vec, u);
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const uint32_t d = LV::digit(order_key<K, FLOATK>(u[j]), 0);
            wave_count(s_hist, d, (full || j * TOPK_HIST_BLOCK + int(threadIdx.x) < int(cnt)) && d < below);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TOPK_BINS; i += TOPK_HIST_BLOCK)
        if (s_hist[i]) atomicAdd(&st->hist[0][i], s_hist[i]);
}

// Block-wide: the bin b of counts h with above(b) < k_rem <= above(b) +
// h[b], above(b) counting the keys in higher bins.  Returns (b, h[b],
// above(b)) to every thread.
__device__ inline void topk_find_bin(const uint32_t *__restrict__ h, uint32_t k_rem, uint32_t (&res)[3])
{
    constexpr int PER = TOPK_BINS / TOPK_BLOCK; // 8 bins per thread, highest first
    constexpr int WAVES = TOPK_BLOCK / WAVE;
    __shared__ uint32_t s_wsum[WAVES];
    __shared__ uint32_t s_res[3];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    uint32_t v[PER], local = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        v[j] = h[TOPK_BINS - 1 - (tid * PER + j)];
        local += v[j];
    }
    uint32_t x = local;
#pragma unroll
    for (int off = 1; off < WAVE; off <<= 1) {
        const uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == WAVE - 1) s_wsum[wave] = x;
    __syncthreads();
    uint32_t above = x - local;
    for (int w = 0; w < wave; ++w) above += s_wsum[w];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        if (above < k_rem && k_rem <= above + v[j]) {
            s_res[0] = uint32_t(TOPK_BINS - 1 - (tid * PER + j));
            s_res[1] = v[j];
            s_res[2] = above;
        }
        above += v[j];
    }
    __syncthreads();
    res[0] = s_res[0];
    res[1] = s_res[1];
    res[2] = s_res[2];
}

// A wave's LDS stage for appends to a global list: lanes append in order,
// and a full stage is flushed with one atomic on the list's counter.
template <typename K> struct WaveStage {
    K *buf;         // TOPK_STAGE keys in LDS
    uint32_t cnt;   // keys staged (wave-uniform)
    K *dst;         // global list
    uint32_t *ctr;  // its length counter
    uint32_t cap;   // its capacity (writes beyond it are dropped)

    __device__ inline void flush()
    {
        if (cnt == 0u) return;
        __builtin_amdgcn_wave_barrier();
        uint32_t base = 0;
        if ((threadIdx.x & (WAVE - 1)) == 0) base = atomicAdd(ctr, cnt);
        base = __shfl(base, 0);
        for (uint32_t i = threadIdx.x & (WAVE - 1); i < cnt; i += WAVE)
            if (base + i < cap) dst[base + i] = buf[i];
        __builtin_amdgcn_wave_barrier();
        cnt = 0u;
    }
    // Write the staged keys at dst[base...] (the slot already reserved).
    __device__ inline void write_at(uint32_t base)
    {
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = threadIdx.x & (WAVE - 1); i < cnt; i += WAVE)
            if (base + i < cap) dst[base + i] = buf[i];
        __builtin_amdgcn_wave_barrier();
        cnt = 0u;
    }
    // Append v from the lanes where `take` holds.
    __device__ inline void append(bool take, K v)
    {
        const uint64_t mk = __ballot(take);
        if (mk == 0) return;
        const uint32_t c = uint32_t(__popcll(mk));
        if (cnt + c > uint32_t(TOPK_STAGE)) flush();
        if (take) buf[cnt + count_below(mk)] = v;
        cnt += c;
    }
};

// Level-0 counts of TOPK_SAMPLE keys spaced n / TOPK_SAMPLE apart (grid:
// TOPK_SAMPLE_BLOCKS workgroups of TOPK_BLOCK threads; n >= TOPK_FILTER_MIN).
template <typename K, int FLOATK>
__global__ __launch_bounds__(TOPK_BLOCK) void k_topk_sample(const K *__restrict__ in, uint32_t n,
                                                            TopkState *__restrict__ st)
{
    using LV = TopkLevels<K>;
    __shared__ uint32_t s_hist[TOPK_BINS];
    for (int i = threadIdx.x; i < TOPK_BINS; i += TOPK_BLOCK) s_hist[i] = 0u;
    const uint32_t stride = n / TOPK_SAMPLE;
    K u[TOPK_SAMPLE_ITEMS];
#pragma unroll
    for (int j = 0; j < TOPK_SAMPLE_ITEMS; ++j) {
        const uint32_t q = (blockIdx.x * uint32_t(TOPK_SAMPLE_ITEMS) + uint32_t(j)) * TOPK_BLOCK + threadIdx.x;
        u[j] = in[size_t(q) * stride + stride / 2];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < TOPK_SAMPLE_ITEMS; ++j) atomicAdd(&s_hist[LV::digit(order_key<K, FLOATK>(u[j]), 0)], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < TOPK_BINS; i += TOPK_BLOCK)
        if (s_hist[i]) atomicAdd(&st->shist[i], s_hist[i]);
}

// One read of the input: the level-0 bin b_est one below the bin of the
// sample's (2 m S / n + 16)-th largest key (S sample keys; every workgroup
// derives the same bin), and every key of a bin >= b_est appended to `kept`
// (raw values, capacity n) and counted into the level-0 counts.
template <typename K, int FLOATK>
__global__ __launch_bounds__(TOPK_BLOCK) void k_topk_filter(const K *__restrict__ in, uint32_t n, int vec,
                                                            TopkState *__restrict__ st, K *__restrict__ kept,
                                                            uint32_t m)
{
    using LV = TopkLevels<K>;
    constexpr int ITEMS = TopkTile<K>::ITEMS, TILE = TopkTile<K>::KEYS;
    constexpr int WAVES = TOPK_BLOCK / WAVE;
    __shared__ uint32_t s_hist[TOPK_BINS];
    __shared__ K s_stage[WAVES][TOPK_STAGE];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    const uint64_t target = min(2ull * uint64_t(m) * TOPK_SAMPLE / uint64_t(n) + 16ull, uint64_t(TOPK_SAMPLE));
    uint32_t sel[3];
    topk_find_bin(st->shist, uint32_t(target), sel);
    const uint32_t b = sel[0] > 0u ? sel[0] - 1u : 0u;
    if (blockIdx.x == 0 && tid == 0) st->b_est = b;
    for (int i = tid; i < TOPK_BINS; i += TOPK_BLOCK) s_hist[i] = 0u;
    __syncthreads();
    WaveStage<K> sk{s_stage[wave], 0u, kept, &st->n_filt, n};
    const uint32_t step = gridDim.x * uint32_t(TILE);
    K nx[ITEMS];
    uint32_t base = blockIdx.x * uint32_t(TILE);
    if (base < n) load_tile_keys<K, TOPK_BLOCK, ITEMS>(in + base, min(uint32_t(TILE), n - base), vec, nx);
    for (; base < n; base += step) {
        const uint32_t cnt = min(uint32_t(TILE), n - base);
        const bool full = vec && cnt == uint32_t(TILE);
        K u[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) u[j] = nx[j];
        if (base + step < n)
            load_tile_keys<K, TOPK_BLOCK, ITEMS>(in + base + step, min(uint32_t(TILE), n - base - step), vec, nx);
        uint32_t d[ITEMS];
        bool any = false;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            d[j] = LV::digit(order_key<K, FLOATK>(u[j]), 0);
            any |= (full || j * TOPK_BLOCK + tid < int(cnt)) && d[j] >= b;
        }
        if (__ballot(any) == 0ull) continue;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool keep = (full || j * TOPK_BLOCK + tid < int(cnt)) && d[j] >= b;
            sk.append(keep, u[j]);
            if (keep) atomicAdd(&s_hist[d[j]], 1u);
        }
    }
    // the waves' remaining stages: one counter atomic per workgroup
    __shared__ uint32_t s_cnt[WAVES], s_base;
    if (lane == 0) s_cnt[wave] = sk.cnt;
    __syncthreads();
    if (tid == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < WAVES; ++w) tot += s_cnt[w];
        s_base = tot ? atomicAdd(sk.ctr, tot) : 0u;
    }
    __syncthreads();
    uint32_t b0 = s_base;
    for (int w = 0; w < wave; ++w) b0 += s_cnt[w];
    sk.write_at(b0);
    for (int i = tid; i < TOPK_BINS; i += TOPK_BLOCK)
        if (s_hist[i]) atomicAdd(&st->hist[0][i], s_hist[i]);
}

// Level `level`: choose the threshold bin b from the level's counts (every
// workgroup computes the same choice; workgroup 0 records it), then classify
// the level's candidates.  Keys above b (and all of b when every key of b is
// needed) are appended to `out` as sort keys; the keys of b become the next
// level's candidates, their next digit counted on the way; lower keys go to
// `rest` as values.  Each wave stages its appends in LDS, so the tiles need
// no workgroup barrier.
template <typename K, int FLOATK>
__global__ __launch_bounds__(TOPK_BLOCK) void k_topk_compact(const K *__restrict__ in, uint32_t n_in, int in_vec,
                                                             const K *__restrict__ cand_src, K *__restrict__ cand_dst,
                                                             int level, TopkState *__restrict__ st,
                                                             K *__restrict__ out, K *__restrict__ rest, uint32_t m,
                                                             int desc, int sampled)
{
    using LV = TopkLevels<K>;
    constexpr int ITEMS = TopkTile<K>::ITEMS, TILE = TopkTile<K>::KEYS;
    constexpr int WAVES = TOPK_BLOCK / WAVE;
    __shared__ uint32_t s_hist[TOPK_BINS];
    __shared__ K s_stage[WAVES][2][TOPK_STAGE];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    const uint32_t done = st->done;
    if (done != 0u && done <= uint32_t(level)) return; // an earlier level took its bin whole
    const uint32_t k_rem = level == 0 ? m : st->sel[level].k_rem;
    const uint64_t prefix = level == 0 ? 0ull : st->sel[level].prefix;
    uint32_t sel[3];
    topk_find_bin(st->hist[level % 3], k_rem, sel);
    const uint32_t b = sel[0], rem = k_rem - sel[2];
    const bool take_bin = sel[1] == rem;
    if (blockIdx.x == 0) {
        if (tid == 0) {
            TopkSel &nx = st->sel[level + 1];
            nx.k_rem = rem;
            nx.bin = b;
            nx.bin_count = sel[1];
            nx.prefix = prefix | (uint64_t(b) << LV::shift(level));
            if (take_bin) st->done = uint32_t(level) + 1u;
            st->n_cand[(level + 2) % 3] = 0u;
        }
        for (int i = tid; i < TOPK_BINS; i += TOPK_BLOCK) st->hist[(level + 2) % 3][i] = 0u;
    }
    const bool last = level + 1 == LV::LEVELS;
    const bool keep_cand = !take_bin && !last;
    const int nlevel = last ? level : level + 1;
    // level 0 reads the filter's kept keys (in cand_src) when they hold the
    // m largest, else the whole input
    const bool filtered = level == 0 && sampled && st->n_filt >= m;
    const uint32_t n = level == 0 ? (filtered ? st->n_filt : n_in) : st->n_cand[level % 3];
    const K *src = level == 0 && !filtered ? in : cand_src;
    const bool vec = level == 0 && !filtered ? in_vec != 0 : true;
    // capacities keep the writes inside the buffers even if the counts were
    // inconsistent (they are not, by construction)
    WaveStage<K> so{s_stage[wave][0], 0u, out, &st->n_out, m};
    WaveStage<K> sc{s_stage[wave][1], 0u, cand_dst, &st->n_cand[(level + 1) % 3], n_in};
    const uint32_t rest_cap = n_in - m;
    if (keep_cand)
        for (int i = tid; i < TOPK_BINS; i += TOPK_BLOCK) s_hist[i] = 0u;
    __syncthreads();
    // the next tile's loads are issued before this tile is classified, so
    // each wave keeps reads in flight while it works
    const uint32_t step = gridDim.x * uint32_t(TILE);
    K nx[ITEMS];
    uint32_t base = blockIdx.x * uint32_t(TILE);
    if (base < n) load_tile_keys<K, TOPK_BLOCK, ITEMS>(src + base, min(uint32_t(TILE), n - base), vec, nx);
    for (; base < n; base += step) {
        const uint32_t cnt = min(uint32_t(TILE), n - base);
        const bool full = vec && cnt == uint32_t(TILE);
        K u[ITEMS];
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) u[j] = nx[j];
        if (base + step < n)
            load_tile_keys<K, TOPK_BLOCK, ITEMS>(src + base + step, min(uint32_t(TILE), n - base - step), vec, nx);
        // most keys lie below the bin: a wave with none at or above it
        // (and no `rest` output to fill) has nothing to append in this tile
        bool any = false;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            if (level == 0) u[j] = order_key<K, FLOATK>(u[j]);
            any |= (full || j * TOPK_BLOCK + tid < int(cnt)) && LV::digit(u[j], level) >= b;
        }
        if (!rest && __ballot(any) == 0ull) continue;
        uint32_t n_rest = 0;
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
            const bool valid = full || j * TOPK_BLOCK + tid < int(cnt);
            const uint32_t d = LV::digit(u[j], level);
            const bool to_out = valid && (d > b || (d == b && take_bin));
            const bool to_cand = valid && keep_cand && d == b;
            so.append(to_out, sort_key(u[j], desc));
            sc.append(to_cand, u[j]);
            if (to_cand) atomicAdd(&s_hist[LV::digit(u[j], nlevel)], 1u);
            n_rest += (rest && valid && d < b) ? 1u : 0u;
        }
        if (rest) {
            // one atomic per wave and tile for the keys below the bin
            uint32_t x = n_rest;
#pragma unroll
            for (int off = 1; off < WAVE; off <<= 1) {
                const uint32_t y = __shfl_up(x, off);
                if (lane >= off) x += y;
            }
            const uint32_t tot = __shfl(x, WAVE - 1);
            uint32_t rbase = 0;
            if (lane == 0 && tot) rbase = atomicAdd(&st->n_rest, tot);
            uint32_t pos = __shfl(rbase, 0) + x - n_rest;
#pragma unroll
            for (int j = 0; j < ITEMS; ++j) {
                const bool valid = full || j * TOPK_BLOCK + tid < int(cnt);
                if (valid && LV::digit(u[j], level) < b) {
                    if (pos < rest_cap) rest[pos] = value_of<K, FLOATK>(u[j]);
                    ++pos;
                }
            }
        }
    }
    // the waves' remaining stages go out with one counter atomic per list and
    // workgroup (one per wave put thousands of atomics on two words)
    __shared__ uint32_t s_cnt[2][WAVES], s_base[2];
    if (lane == 0) {
        s_cnt[0][wave] = so.cnt;
        s_cnt[1][wave] = sc.cnt;
    }
    __syncthreads();
    if (tid < 2) {
        uint32_t tot = 0;
        for (int w = 0; w < WAVES; ++w) tot += s_cnt[tid][w];
        s_base[tid] = tot ? atomicAdd(tid == 0 ? so.ctr : sc.ctr, tot) : 0u;
    }
    __syncthreads();
    uint32_t b0 = s_base[0], b1 = s_base[1];
    for (int w = 0; w < wave; ++w) {
        b0 += s_cnt[0][w];
        b1 += s_cnt[1][w];
    }
    so.write_at(b0);
    sc.write_at(b1);
    if (keep_cand) {
        __syncthreads();
        uint32_t *hn = st->hist[(level + 1) % 3];
        for (int i = tid; i < TOPK_BINS; i += TOPK_BLOCK)
            if (s_hist[i]) atomicAdd(&hn[i], s_hist[i]);
    }
}

// The threshold key T and its counts after the last level.
template <typename K> struct TopkTail {
    K T;
    uint32_t need;    // copies of T that close the output
    uint32_t surplus; // copies of T that close rest
};
template <typename K>
__device__ inline TopkTail<K> topk_tail(const TopkState *__restrict__ st, uint32_t m, uint32_t n)
{
    TopkTail<K> t{K(0), 0u, 0u};
    if (st->done != 0u) return t;
    const TopkSel &s = st->sel[TopkLevels<K>::LEVELS];
    t.T = K(s.prefix);
    t.need = min(s.k_rem, m);
    t.surplus = min(s.bin_count - t.need, n - m);
    return t;
}

// Copies of T at the ends of the output (positions [m - need, m)) and rest.
template <typename K, int FLOATK>
__global__ void k_topk_finish(const TopkState *__restrict__ st, K *__restrict__ out, K *__restrict__ rest,
                              uint32_t m, uint32_t n, int desc)
{
    const TopkTail<K> t = topk_tail<K>(st, m, n);
    const uint32_t stride = gridDim.x * blockDim.x;
    const K s = sort_key(t.T, desc);
    if (out)
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.need; i += stride) out[m - t.need + i] = s;
    if (rest) {
        const K v = value_of<K, FLOATK>(t.T);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < t.surplus; i += stride)
            rest[n - m - t.surplus + i] = v;
    }
}

// Sort keys back to values: ~s for descending order, then the inverse float
// transform.
template <typename K, int FLOATK> __device__ inline K topk_value(K s, int desc)
{
    return value_of<K, FLOATK>(desc ? K(~s) : s);
}

// One workgroup: LSD radix sort in LDS of up to BLOCK * TOPK_SORT_ITEMS sort
// keys (the passes over the digit positions that vary; tile_rank_stage per
// pass), then the first m written to dst as values.  Sources: mode 0 reads
// values src[0, len) (the whole input of a small top-k), mode 1 the sort keys
// src[0, m - need) appended by the compaction plus need copies of the
// threshold key (from the state).  Equal sort keys are equal values, so the
// order among them is immaterial.  src may be dst: every key is loaded
// before the first barrier and written after the last.
template <typename K, int FLOATK, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_topk_radix_sort(const K *__restrict__ src, uint32_t len,
                                                          K *__restrict__ dst, uint32_t m, int mode,
                                                          const TopkState *__restrict__ st, uint32_t n, int desc)
{
    constexpr int ITEMS = TOPK_SORT_ITEMS;
    constexpr int WAVES = BLOCK / WAVE;
    constexpr int WAVE_TILE = WAVE * ITEMS;
    constexpr int TILE = BLOCK * ITEMS;
    constexpr int P = int(sizeof(K));
    // LDS-OR match ranking where its masks fit beside the keys (u64 keys in
    // 1024-thread tiles would need 176 KiB): faster than ballot ranking
    constexpr bool LDSOR = size_t(TILE) * sizeof(K) + size_t(WAVES) * RADIX * 12 <= 144 * 1024;
    __shared__ uint32_t s_cnt[WAVES][RADIX];
    __shared__ uint64_t s_match[LDSOR ? WAVES : 1][LDSOR ? RADIX : 1];
    __shared__ uint32_t s_scan[RADIX / WAVE];
    __shared__ uint32_t s_start[RADIX];
    __shared__ K s_keys[TILE];
    __shared__ K s_or[WAVES], s_nor[WAVES];
    const int tid = threadIdx.x, lane = tid & (WAVE - 1), wave = tid >> 6;
    uint32_t have = len;
    K fill = K(0);
    if (mode == 1) {
        const TopkTail<K> t = topk_tail<K>(st, m, n);
        fill = sort_key(t.T, desc);
        have = m - t.need;
        len = m;
    }
    const int cnt = int(min(len, uint32_t(TILE)));
    const int wave_valid = cnt > wave * WAVE_TILE ? min(WAVE_TILE, cnt - wave * WAVE_TILE) : 0;
    K v[ITEMS];
    K vor = 0, vnor = 0;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const int o = i * WAVE + lane;
        const uint32_t q = uint32_t(wave * WAVE_TILE + o);
        v[i] = ~K(0);
        if (o < wave_valid) {
            K x = fill;
            if (q < have) x = mode == 0 ? sort_key(order_key<K, FLOATK>(src[q]), desc) : src[q];
            v[i] = x;
            vor |= x;
            vnor |= ~x;
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
    bool staged = false;
    for (int pass = 0; pass < P; ++pass) {
        if (((varying >> (8 * pass)) & K(0xFF)) == 0) continue;
        if (staged) {
#pragma unroll
            for (int i = 0; i < ITEMS; ++i) {
                const int o = i * WAVE + lane;
                if (o < wave_valid) v[i] = s_keys[wave * WAVE_TILE + o];
            }
            __syncthreads();
        }
        if constexpr (LDSOR)
            tile_rank_stage_ldsor<K, BLOCK, ITEMS>(v, wave_valid, 8 * pass, s_cnt, s_match, s_scan, s_start, s_keys);
        else
            tile_rank_stage<K, BLOCK, ITEMS>(v, wave_valid, 8 * pass, s_cnt, s_scan, s_start, s_keys);
        staged = true;
    }
    if (!staged) { // every key equal: any order is sorted
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const int o = i * WAVE + lane;
            if (o < wave_valid) s_keys[wave * WAVE_TILE + o] = v[i];
        }
        __syncthreads();
    }
    const uint32_t out = min(m, uint32_t(cnt));
    for (uint32_t j = tid; j < out; j += BLOCK) dst[j] = topk_value<K, FLOATK>(s_keys[j], desc);
}

// Up to TOPK_RANK_SORT keys: one thread per key counts the keys that order
// before it ((key, index) pairs, so ranks are distinct) against the whole
// set in LDS and writes itself to its rank.  Same sources and modes as
// k_topk_radix_sort; blockDim.x is a power of two >= len.
template <typename K, int FLOATK>
__global__ __launch_bounds__(TOPK_RANK_SORT) void k_topk_rank_sort(const K *__restrict__ src, uint32_t len,
                                                                   K *__restrict__ dst, uint32_t m, int mode,
                                                                   const TopkState *__restrict__ st, uint32_t n,
                                                                   int desc)
{
    __shared__ K s[TOPK_RANK_SORT];
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    uint32_t have = len;
    K fill = K(0);
    if (mode == 1) {
        const TopkTail<K> t = topk_tail<K>(st, m, n);
        fill = sort_key(t.T, desc);
        have = m - t.need;
        len = m;
    }
    K x = src[min(tid, have > 0u ? have - 1u : 0u)];
    x = mode == 0 ? sort_key(order_key<K, FLOATK>(x), desc) : x;
    x = tid < have ? x : (tid < len ? fill : K(~K(0))); // padding sorts last
    s[tid] = x;
    __syncthreads();
    uint32_t rank = 0;
    for (uint32_t i = 0; i < nt; i += 4) {
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e) {
            const K y = s[i + e];
            rank += (y < x || (y == x && i + e < tid)) ? 1u : 0u;
        }
    }
    if (tid < len && rank < m) dst[rank] = topk_value<K, FLOATK>(x, desc);
}

template <typename K, int FLOATK>
__global__ void k_topk_values(K *__restrict__ keys, uint64_t len, int desc)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += stride)
        keys[i] = topk_value<K, FLOATK>(keys[i], desc);
}

} // namespace hpcs
