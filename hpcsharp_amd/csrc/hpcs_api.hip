// hpcs_api.hip -- host engine and C ABI of libhpcsort.so.
//
// The engine owns, per context, a device scratch pool, a stream, tuning
// options and a tuning table.  A sort is a fixed sequence of stream-ordered
// launches chosen from (dtype, n): for large arrays fill, upsweep, plan and
// per pass tile counts, three scan kernels and the scatter (reduce-then-
// scan); for small ones fill, histogram, plan and one onesweep kernel per
// pass; then a copy-back.  No host synchronisation happens inside a
// device-resident call, so the sequence can be captured in a HIP graph.
#include "hpcs_kernels.hpp"
#include "hpcs_shapes.hpp"
#include "hpcs_digit11.hpp"
#include "hpcs_merge.hpp"
#include "hpcs_msd.hpp"
#include "hpcs_segments.hpp"
#include "hpcs_hybrid.hpp"
#include "hpcs_topk.hpp"
#include "hpcs_tuning.hpp"
#include "../../include/hpcsort.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

using namespace hpcs;

#define HPCS_API extern "C" __attribute__((visibility("default")))

// Split compilation: hpcsharp_amd/build.py compiles this file once per part
// (-DHPCS_PART=k, in parallel), each object holding one group of entry points
// and the kernels they instantiate; without HPCS_PART one object holds all.
#ifdef HPCS_PART
#define HPCS_IN_PART(k) (HPCS_PART == (k))
#else
#define HPCS_IN_PART(k) 1
#endif

namespace hpcs_state {
// the calling thread's last error message: one object for all parts
inline thread_local std::string g_last_error;
} // namespace hpcs_state

namespace {

using hpcs_state::g_last_error;

struct Err {
    int32_t code;
    std::string msg;
};

[[noreturn]] void fail(int32_t code, const std::string &msg) { throw Err{code, msg}; }

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) fail(HPCS_E_OOM, std::string(what) + ": out of device memory");
        fail(HPCS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

constexpr int PASS_ONESWEEP = 0;
constexpr int PASS_RTS = 1;
constexpr int PASS_AUTO = 2;
constexpr int PASS_BLOCK = 3; // one workgroup sorts the whole array in LDS
constexpr int PASS_MSD = 4;   // one MSD split, then one workgroup per bucket (keys without values)
constexpr int PASS_MERGE = 5; // runs sorted by one workgroup each in LDS, then stable merge passes
constexpr int PASS_HYBRID = 6; // 4-byte unsigned keys without values: top two bytes MSD-first, then LDS buckets
// Largest sort of the MSD-first form: its 65536 buckets then hold 4608 keys on
// average for uniform keys, within the LDS sort's HYB_CAP (5120) with margin;
// larger arrays take the reduce-then-scan passes.
constexpr int64_t HYB_MAX_KEYS = (int64_t(1) << 28) + (int64_t(1) << 25);
// largest arrays the block-sort + merge form serves (beyond, the passes over
// the whole array outnumber the radix passes)
constexpr int64_t MERGE_SORT_MAX = int64_t(1) << 20;
// largest arrays the MSD split serves (beyond, most buckets outgrow a tile)
constexpr int64_t MSD_MAX4 = int64_t(1) << 22, MSD_MAX8 = int64_t(1) << 21;
// largest arrays the one-workgroup sort takes (LDS: keys, values, counters)
constexpr int64_t BLOCK_SORT_MAX4 = 1024 * 16; // 4-byte keys
constexpr int64_t BLOCK_SORT_MAX8 = 1024 * 8;  // 8-byte keys


// Tile shapes: TILE_SHAPES (hpcs_shapes.hpp).  with_shape runs f with the
// row as a compile-time constant; every row that serves the key type has a
// branch, and any other index fails (no default instantiation runs), so the
// counting and scattering launchers of one sort always use one tile.
template <int KB, bool V, int S = 0, typename F>
void with_shape(int shape, F &&f)
{
    if constexpr (S < NUM_TILE_SHAPES) {
        if (shape == S) {
            if constexpr (shape_serves(S, KB, V)) {
                f(std::integral_constant<int, S>{});
                return;
            } else {
                fail(HPCS_E_UNSUPPORTED, "tile shape " + std::to_string(shape) + " does not serve this key type");
            }
        }
        with_shape<KB, V, S + 1>(shape, std::forward<F>(f));
    } else {
        fail(HPCS_E_UNSUPPORTED, "unknown tile shape " + std::to_string(shape));
    }
}

bool read_file(const char *path, std::string &text)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[4096];
    size_t got;
    text.clear();
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, got);
    std::fclose(f);
    return true;
}
constexpr int SORT_TILE = shape_tile_keys(0);
// k_histogram's shape per key size: threads per workgroup, interleaved
// copies of its LDS table (64 KiB per workgroup either way), 1 KiB chunks per
// wave in flight.  Measured at 2^28 keys on MI355X against the earlier 256
// threads x 4 copies (2 for 8-byte keys) x 4 chunks: uniform u32 230 -> 172
// us, sorted 188 -> 160 us; uniform u64 528 -> 414 us.
template <typename K> constexpr int hist_block() { return sizeof(K) == 4 ? 768 : 1024; }
template <typename K> constexpr int hist_copies() { return sizeof(K) == 4 ? 16 : 8; }
template <typename K> constexpr int hist_unroll() { return sizeof(K) == 4 ? 2 : 8; }
constexpr size_t ALIGN = 256;

size_t align_up(size_t x, size_t a = ALIGN) { return (x + a - 1) / a * a; }

} // namespace

// ---------------------------------------------------------------------------
// Scratch of hpcs_dev_sort_radix_u32_noalloc: a device array in the code
// object, one copy per GPU, held until the process ends.  Host code never
// passes `scratch` itself to a kernel or a copy (on the host the name is a
// stand-in, not a device address): each context resolves the array's device
// address on its GPU once, at creation, with hipGetSymbolAddress, and the
// entry point passes that address.  Every call on one GPU shares the array,
// so overlapping calls on one GPU race; the sort clears what must start at
// zero itself.  Largest sort: 2^22 keys (the layout of a 2^22-key sort with
// 2048-key tiles, the smallest tile, is 25.3 MiB).
// ---------------------------------------------------------------------------
constexpr int64_t NOALLOC_MAX_KEYS = int64_t(1) << 22;
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(26) << 20;
#if HPCS_IN_PART(0) // the array and its address lookup live in the part that creates contexts
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
#endif

struct hpcs_ctx {
    int device = 0;
    void *noalloc_scratch = nullptr; // device address of `scratch` on this context's GPU
    std::string noalloc_error;       // why it is null
    int num_cus = 256;
    hipStream_t stream = nullptr;
    std::mutex mu;
    void *scratch = nullptr;
    size_t scratch_cap = 0;
    void *io = nullptr;
    size_t io_cap = 0;
    // record and item sorts: [keys][indices][copy of the records], grown on
    // demand like the scratch
    void *work = nullptr;
    size_t work_cap = 0;
    int rank_mode = RANK_LDSOR;  // tuning knob, env HPCS_RANK=ballot|ldsor|mix2|mix4
    int tile_shape = 0;          // tuning knob, env HPCS_TILE_SHAPE: a row of TILE_SHAPES
    int fallback_rounds = FALLBACK_ROUNDS;
    int lookback_window = 32;    // u32 keys: 16, 32 or 64 earlier tiles per look-back round
    int pass_algo = PASS_AUTO;    // PASS_ONESWEEP, PASS_RTS (reduce-then-scan) or PASS_AUTO (by size)
    size_t scratch_limit = 0;     // bytes of scratch a sort may use (0: no limit); keys-only sorts that
                                  // need more sort each half and merge (HPCS_OPT_SCRATCH_LIMIT)
    int digit_bytes = 2;          // reduce-then-scan: tile counts from next-digit bytes written by the
                                  // scatter: 0 off, 1 on, 2 by type (on for 8-byte keys with values)
    // tuning table (hpcs_ctx_load_tuning / HPCS_TUNE_FILE); options set
    // explicitly with hpcs_ctx_set_option or the environment take precedence
    int digit_bits = 8;           // 8, or 11 for onesweep sorts of keys without values
    // batched segments: ranking of the large segments' scatter, largest
    // segment sorted in LDS (HPCS_OPT_SEG_RANK, HPCS_OPT_SEG_SMALL)
    int seg_rank = RANK_MIX2;
    int seg_small = SEG_SMALL;
    int seg_shape = 0;
    int64_t seg_msd = SEG_MSD_DEFAULT; // segments longer than this take the MSD-first form (0: none)
    std::vector<TuneEntry> tune;
    bool set_algo = false, set_shape = false, set_rank = false, set_bits = false;
    bool set_seg_rank = false, set_seg_small = false, set_seg_shape = false;
    // kernel timing
    struct Pending {
        const char *name;
        hipEvent_t a, b;
    };
    bool profile = false;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> ev_pool;
    std::map<std::string, std::pair<int64_t, double>> stats;
    // graph cache (HPCS_OPT_GRAPHS): a device-resident call that repeats
    // (same entry point, pointers, size and stream) is replayed from a HIP
    // graph of its launch sequence
    struct GraphEntry {
        int entry;
        const void *p0, *p1;
        int64_t n, aux;
        hipStream_t st;
        void *scratch, *io; // the context buffers when the call last ran directly
        size_t scratch_cap, io_cap;
        hipGraphExec_t exec; // null until the call repeats
        uint64_t last_use;
    };
    int graphs = 0;
    std::vector<GraphEntry> graph_cache;
    uint64_t graph_clock = 0;
    int64_t graph_captures = 0, graph_replays = 0;
};

namespace {

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        hip_check(hipGetDevice(&prev), "hipGetDevice");
        if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard()
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

// Grow a cached device buffer.  Work queued on any stream may still use the
// old buffer, so the device is drained before it is released.
void *ensure(void *&buf, size_t &cap, size_t need)
{
    if (need <= cap && buf) return buf;
    if (buf) {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        hip_check(hipFree(buf), "hipFree");
        buf = nullptr;
        cap = 0;
    }
    size_t want = align_up(std::max(need, size_t(1) << 20), size_t(2) << 20);
    hip_check(hipMalloc(&buf, want), "hipMalloc");
    cap = want;
    return buf;
}

// Scratch layout of one sort:
//   [hist P x 256 u32 | DevPlan | look-back region 0]  zero-filled first
//   [look-back region 1] [16-bit tile counts] [keys B] [values B]
//   [next-pass digit bytes, n]
struct Layout {
    size_t off_hist, off_plan, off_dtot, off_lb0, zero_bytes, off_lb1, off_cnt, off_keysB, off_valsB, off_dig, total;
    int64_t tiles;
};

static_assert(ORSLOTS * 128 <= MAX_PASSES * RADIX * sizeof(uint32_t), "OR slots live in the histogram area");

Layout make_layout(int64_t n, size_t key_size, bool vals, int tile_keys = SORT_TILE, bool dig = true)
{
    Layout L{};
    L.tiles = (n + tile_keys - 1) / tile_keys;
    const size_t lb_bytes = size_t(L.tiles) * RADIX * sizeof(uint32_t);
    L.off_hist = 0;
    L.off_plan = align_up(MAX_PASSES * RADIX * sizeof(uint32_t));
    L.off_dtot = align_up(L.off_plan + sizeof(DevPlan));
    L.off_lb0 = align_up(L.off_dtot + size_t(MAX_PASSES) * RADIX * LINE_WORDS * sizeof(uint32_t));
    L.zero_bytes = L.off_lb0 + lb_bytes;
    L.off_lb1 = align_up(L.zero_bytes);
    L.off_cnt = align_up(L.off_lb1 + lb_bytes); // 16-bit tile counts of the reduce-then-scan passes
    L.off_keysB = align_up(L.off_cnt + lb_bytes / 2);
    L.off_valsB = align_up(L.off_keysB + size_t(n) * key_size);
    L.off_dig = align_up(L.off_valsB + (vals ? size_t(n) * sizeof(uint32_t) : 0));
    L.total = align_up(L.off_dig + (dig ? size_t(n) : 0));
    return L;
}

// The reduce-then-scan scatter writes the next pass's digit bytes for its
// tile counts (HPCS_OPT_DIGIT_BYTES).  Measured at 2^28 on MI355X: the byte
// stores cost the scatter about what the byte counts save for u64 keys, more
// for u32 keys; with u32 values beside u64 keys the sort gains 5%.  Sorts
// without them leave the layout's n-byte digit region out.
bool digit_bytes_on(const hpcs_ctx *c, size_t key_size, bool vals)
{
    return c->digit_bytes == 1 || (c->digit_bytes == 2 && key_size == 8 && vals);
}

hipEvent_t take_event(hpcs_ctx *c)
{
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    return e;
}

// Wait for recorded launch events and fold them into the statistics.
void harvest(hpcs_ctx *c)
{
    for (auto &p : c->pending) {
        hip_check(hipEventSynchronize(p.b), "hipEventSynchronize");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        auto &s = c->stats[p.name];
        s.first += 1;
        s.second += ms;
        c->ev_pool.push_back(p.a);
        c->ev_pool.push_back(p.b);
    }
    c->pending.clear();
}

// Run `launch` (which enqueues one kernel on st), bracketed by events when
// profiling is on.
template <typename F>
void timed(hpcs_ctx *c, hipStream_t st, const char *name, F &&launch)
{
    if (!c->profile) {
        launch();
        hip_check(hipGetLastError(), name);
        return;
    }
    if (c->pending.size() >= 4096) harvest(c);
    hipEvent_t a = take_event(c), b = take_event(c);
    hip_check(hipEventRecord(a, st), "hipEventRecord");
    launch();
    hip_check(hipGetLastError(), name);
    hip_check(hipEventRecord(b, st), "hipEventRecord");
    c->pending.push_back({name, a, b});
}

int grid_for(uint64_t items, int block, int max_blocks)
{
    uint64_t g = (items + block - 1) / block;
    if (g < 1) g = 1;
    if (g > uint64_t(max_blocks)) g = max_blocks;
    return int(g);
}

void launch_fill(hpcs_ctx *c, void *p, uint64_t bytes, hipStream_t st)
{
    const uint64_t words = bytes / 4;
    if (!words) return;
    timed(c, st, "fill", [&] {
        hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(words, 256, c->num_cus * 8)), dim3(256), 0, st,
                           static_cast<uint32_t *>(p), words, 0u);
    });
}

template <typename K, int FLOATK>
void launch_histogram(hpcs_ctx *c, const K *d_keys, uint64_t n, uint32_t *d_hist, hipStream_t st)
{
    constexpr int VEC = 16 / int(sizeof(K));
    const uintptr_t addr = reinterpret_cast<uintptr_t>(d_keys);
    if (addr % sizeof(K)) fail(HPCS_E_INVALID_ARG, "keys pointer is not aligned to the key size");
    const uint32_t head = uint32_t(((16 - (addr & 15)) & 15) / sizeof(K));
    const uint64_t chunks = (n > head ? n - head : 0) / (WAVE * VEC);
    constexpr int BLOCK = hist_block<K>(), COPIES = hist_copies<K>(), UNROLL = hist_unroll<K>();
    const int waves_per_block = BLOCK / WAVE;
    // at most the resident workgroups (a grid-stride loop: a second round of
    // workgroups would run on a partly idle chip)
    static const int occ = [] {
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_histogram<K, FLOATK, BLOCK, COPIES, UNROLL>, BLOCK, 0) !=
                hipSuccess ||
            o < 1)
            o = 1;
        return std::min(o, 8);
    }();
    const int grid = grid_for(chunks, waves_per_block * UNROLL, c->num_cus * occ);
    timed(c, st, "histogram", [&] {
        hipLaunchKernelGGL((k_histogram<K, FLOATK, BLOCK, COPIES, UNROLL>), dim3(grid), dim3(BLOCK), 0, st, d_keys, n,
                           head, d_hist);
    });
}

template <typename K, bool HAS_VALS, int FLOATK, int BLOCK, int ITEMS, int W>
void launch_onesweep_w(int rank_mode, int fb, dim3 grid, hipStream_t st, K *a, K *b, uint32_t *va, uint32_t *vb,
                       uint32_t n, int pass, DevPlan *plan, uint32_t *lb, uint32_t *next, uint8_t *dig)
{
    // 1024-thread tiles rank by ballot only (their LDS holds no match masks)
    if (BLOCK >= 1024 || rank_mode == RANK_BALLOT)
        hipLaunchKernelGGL((k_onesweep<K, HAS_VALS, FLOATK, BLOCK, ITEMS, RANK_BALLOT, W>), grid, dim3(BLOCK), 0, st,
                           a, b, va, vb, n, pass, plan, lb, next, fb, dig);
    else if constexpr (BLOCK < 1024) {
        if (rank_mode == RANK_MIX2)
            hipLaunchKernelGGL((k_onesweep<K, HAS_VALS, FLOATK, BLOCK, ITEMS, RANK_MIX2, W>), grid, dim3(BLOCK), 0,
                               st, a, b, va, vb, n, pass, plan, lb, next, fb, dig);
        else if (rank_mode == RANK_MIX4)
            hipLaunchKernelGGL((k_onesweep<K, HAS_VALS, FLOATK, BLOCK, ITEMS, RANK_MIX4, W>), grid, dim3(BLOCK), 0,
                               st, a, b, va, vb, n, pass, plan, lb, next, fb, dig);
        else
            hipLaunchKernelGGL((k_onesweep<K, HAS_VALS, FLOATK, BLOCK, ITEMS, RANK_LDSOR, W>), grid, dim3(BLOCK), 0,
                               st, a, b, va, vb, n, pass, plan, lb, next, fb, dig);
    }
}

template <typename K, bool HAS_VALS, int FLOATK, int BLOCK, int ITEMS>
void launch_onesweep_shape(int rank_mode, int window, int fb, dim3 grid, hipStream_t st, K *a, K *b, uint32_t *va,
                           uint32_t *vb, uint32_t n, int pass, DevPlan *plan, uint32_t *lb, uint32_t *next,
                           uint8_t *dig)
{
    if (window == 0)
        return launch_onesweep_w<K, HAS_VALS, FLOATK, BLOCK, ITEMS, 0>(rank_mode, fb, grid, st, a, b, va, vb, n,
                                                                       pass, plan, lb, next, dig);
    // windows 16 and 64 serve the default 256x16 tiles (others take 32)
    if constexpr (sizeof(K) == 4 && !HAS_VALS && BLOCK == 256 && ITEMS == 16) {
        if (window <= 16)
            return launch_onesweep_w<K, HAS_VALS, FLOATK, BLOCK, ITEMS, 16>(rank_mode, fb, grid, st, a, b, va, vb, n,
                                                                            pass, plan, lb, next, dig);
        if (window >= 64)
            return launch_onesweep_w<K, HAS_VALS, FLOATK, BLOCK, ITEMS, 64>(rank_mode, fb, grid, st, a, b, va, vb, n,
                                                                            pass, plan, lb, next, dig);
    }
    launch_onesweep_w<K, HAS_VALS, FLOATK, BLOCK, ITEMS, 32>(rank_mode, fb, grid, st, a, b, va, vb, n, pass, plan,
                                                             lb, next, dig);
}

// One pass: onesweep with look-back (window > 0) or the prescanned scatter
// (window 0; lb holds the tile offsets).
template <typename K, bool HAS_VALS, int FLOATK>
void launch_onesweep(int rank_mode, int shape, int window, int fb, dim3 grid, hipStream_t st, K *a, K *b,
                     uint32_t *va, uint32_t *vb, uint32_t n, int pass, DevPlan *plan, uint32_t *lb, uint32_t *next,
                     uint8_t *dig)
{
    with_shape<int(sizeof(K)), HAS_VALS>(shape, [&](auto sc) {
        constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
        launch_onesweep_shape<K, HAS_VALS, FLOATK, T.block, T.items>(rank_mode, window, fb, grid, st, a, b, va, vb,
                                                                     n, pass, plan, lb, next, dig);
    });
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Per-tile digit counts of one pass (grid = one workgroup per tile).
template <typename K, bool HAS_VALS, int FLOATK>
void launch_tile_counts(int shape, dim3 grid, hipStream_t st, const K *a, const K *b, uint32_t n, int pass,
                        const DevPlan *plan, uint16_t *counts, const uint8_t *dig)
{
    const int vec = aligned16(a) && aligned16(b);
    with_shape<int(sizeof(K)), HAS_VALS>(shape, [&](auto sc) {
        constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
        hipLaunchKernelGGL((k_tile_counts<K, FLOATK, T.block, T.items>), grid, dim3(T.block), 0, st, a, b, n, pass,
                           plan, counts, vec, dig);
    });
}

// The first read of a reduce-then-scan sort (grid = one workgroup per tile).
template <typename K, bool HAS_VALS, int FLOATK>
void launch_upsweep(int shape, dim3 grid, hipStream_t st, const K *a, uint32_t n, uint16_t *counts, K *orslots)
{
    const int vec = aligned16(a);
    with_shape<int(sizeof(K)), HAS_VALS>(shape, [&](auto sc) {
        constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
        hipLaunchKernelGGL((k_upsweep<K, FLOATK, T.block, T.items>), grid, dim3(T.block), 0, st, a, n, counts,
                           orslots, vec);
    });
}

// The device-resident sort: keys (and optional uint32 values) in place.
struct SortParams {
    int algo, shape, rank, bits;
};

// The pass algorithm, tile shape and ranking for a sort of n keys: options
// set explicitly win, then the tuning table's row for the smallest size class
// that holds n (the largest class beyond it), then the built-in defaults.
SortParams resolve_params(const hpcs_ctx *c, int key_bytes, bool vals, int64_t n)
{
    SortParams p{c->pass_algo, shape_serves(c->tile_shape, key_bytes, vals) ? c->tile_shape : 0, c->rank_mode,
                 c->digit_bits};
    const TuneEntry *best = nullptr, *largest = nullptr;
    for (const TuneEntry &e : c->tune) {
        if (e.key_bytes != key_bytes || e.vals != int(vals)) continue;
        if (!largest || e.max_n > largest->max_n) largest = &e;
        if (e.max_n >= n && (!best || e.max_n < best->max_n)) best = &e;
    }
    if (!best) best = largest;
    if (best) {
        if (!c->set_algo) p.algo = best->algo;
        if (!c->set_shape && shape_serves(best->shape, key_bytes, vals)) p.shape = best->shape;
        if (!c->set_rank) p.rank = best->rank;
        if (!c->set_bits) p.bits = best->bits;
    }
    // Below these sizes the single-kernel onesweep pass beats reduce-then-
    // scan's five kernels per pass (launch-bound), measured on MI355X.
    const int64_t block_max = key_bytes == 4 ? BLOCK_SORT_MAX4 : BLOCK_SORT_MAX8;
    if (p.algo == PASS_BLOCK && n > block_max) p.algo = PASS_AUTO;
    if (p.algo == PASS_MSD && (vals || n > (key_bytes == 4 ? MSD_MAX4 : MSD_MAX8))) p.algo = PASS_AUTO;
    if (p.algo == PASS_MERGE && n > MERGE_SORT_MAX) p.algo = PASS_AUTO;
    // the MSD-first form sorts 4-byte keys without values (run_sort sends
    // float and signed keys to the reduce-then-scan passes)
    if (p.algo == PASS_HYBRID && (key_bytes != 4 || vals || n <= BLOCK_SORT_MAX4 || n > HYB_MAX_KEYS))
        p.algo = PASS_AUTO;
    // by size, measured on MI355X: one workgroup up to 16384 4-byte / 8192
    // 8-byte keys (with LDS-OR ranking it beats onesweep passes there: u32
    // 16384 keys 42.2 against 48.4 us)
    if (p.algo == PASS_AUTO)
        p.algo = n <= (key_bytes == 4 ? BLOCK_SORT_MAX4 : BLOCK_SORT_MAX8) ? PASS_BLOCK
                 : n < (key_bytes == 4 ? (int64_t(1) << 19) : (int64_t(1) << 21)) ? PASS_ONESWEEP
                                                                                   : PASS_RTS;
    // 11-bit digits exist for onesweep passes over keys without values (their
    // look-back words carry 30-bit counts); other sorts keep 8-bit digits
    if (p.bits == 11 && (vals || p.algo != PASS_ONESWEEP || n >= LOOKBACK_MAX_KEYS)) p.bits = 8;
    if (p.bits == 11) {
        p.shape = 0;
        p.rank = RANK_BALLOT;
    }
    if (TILE_SHAPES[p.shape].ballot_only) p.rank = RANK_BALLOT;
    return p;
}

// The one-workgroup sort (n <= BLOCK_SORT_MAX4 / BLOCK_SORT_MAX8): no scratch.
template <typename K, bool HAS_VALS, int FLOATK>
void run_sort_block(hpcs_ctx *c, K *d_keys, uint32_t *d_vals, int64_t n, hipStream_t st)
{
    constexpr int ITEMS = sizeof(K) == 4 ? 16 : 8;
    timed(c, st, "block_sort", [&] {
        if (n <= 256 * ITEMS)
            hipLaunchKernelGGL((k_sort_block<K, FLOATK, HAS_VALS, 256, ITEMS>), dim3(1), dim3(256), 0, st, d_keys,
                               d_vals, d_keys, d_vals, uint32_t(n));
        else
            hipLaunchKernelGGL((k_sort_block<K, FLOATK, HAS_VALS, 1024, ITEMS>), dim3(1), dim3(1024), 0, st, d_keys,
                               d_vals, d_keys, d_vals, uint32_t(n));
    });
}

// The block-sort + merge form (n <= MERGE_SORT_MAX): k_sort_block with one
// workgroup per run of MERGE_RUN keys (a 256-thread tile) sorts every run in
// LDS, then k_merge_runs merges pairs of runs until one is left, ping-pong
// between the caller's buffer and n keys (and values) of scratch; the block
// sort writes where the passes must start so that the last pass lands in the
// caller's buffer.  1 + ceil(log2(runs)) launches, no counting pass, no plan.
template <typename K> constexpr int merge_run_items() { return sizeof(K) == 4 ? 16 : 8; }
template <typename K> constexpr int merge_run_keys() { return 256 * merge_run_items<K>(); }
static_assert(256 * 8 % MERGE_TILE == 0, "a run is a multiple of the merge tile");

size_t merge_sort_scratch_bytes(int64_t n, size_t key_size, bool vals)
{
    return align_up(size_t(n) * key_size) + (vals ? align_up(size_t(n) * sizeof(uint32_t)) : 0);
}

template <typename K, bool HAS_VALS, int FLOATK>
void run_sort_merge(hpcs_ctx *c, K *d_keys, uint32_t *d_vals, int64_t n, hipStream_t st)
{
    constexpr int RUN = merge_run_keys<K>();
    const int64_t runs = (n + RUN - 1) / RUN;
    int merges = 0;
    while ((int64_t(RUN) << merges) < n) ++merges;
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, merge_sort_scratch_bytes(n, sizeof(K), HAS_VALS)));
    K *kb = reinterpret_cast<K *>(base);
    uint32_t *vb = HAS_VALS ? reinterpret_cast<uint32_t *>(base + align_up(size_t(n) * sizeof(K))) : nullptr;
    // after an odd number of merge passes the runs must start in scratch
    K *ks = (merges & 1) ? kb : d_keys, *kd = (merges & 1) ? d_keys : kb;
    uint32_t *vs = (merges & 1) ? vb : d_vals, *vd = (merges & 1) ? d_vals : vb;
    timed(c, st, "block_sort", [&] {
        hipLaunchKernelGGL((k_sort_block<K, FLOATK, HAS_VALS, 256, merge_run_items<K>()>), dim3(unsigned(runs)),
                           dim3(256), 0, st, d_keys, d_vals, ks, vs, uint32_t(n));
    });
    const unsigned grid = unsigned((n + MERGE_TILE - 1) / MERGE_TILE);
    for (int64_t run = RUN; run < n; run *= 2) {
        timed(c, st, "merge", [&] {
            hipLaunchKernelGGL((k_merge_runs<K, HAS_VALS, FLOATK>), dim3(grid), dim3(MERGE_BLOCK), 0, st, ks, vs, kd,
                               vd, uint64_t(n), uint64_t(run));
        });
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
}

// Scratch of an 11-bit sort:
//   [hist passes x 2048 u32 | DevPlan11 | look-back region 0]  zero-filled
//   [look-back region 1] [keys B]
struct Layout11 {
    size_t off_plan, off_lb0, zero_bytes, off_lb1, off_keysB, total;
    int64_t tiles;
};

Layout11 make_layout11(int64_t n, size_t key_size)
{
    Layout11 L{};
    L.tiles = (n + D11_TILE - 1) / D11_TILE;
    const size_t lb_bytes = size_t(L.tiles) * D11_RADIX * sizeof(uint32_t);
    L.off_plan = align_up(D11_MAX_PASSES * D11_RADIX * sizeof(uint32_t));
    L.off_lb0 = align_up(L.off_plan + sizeof(DevPlan11));
    L.zero_bytes = L.off_lb0 + lb_bytes;
    L.off_lb1 = align_up(L.zero_bytes);
    L.off_keysB = align_up(L.off_lb1 + lb_bytes);
    L.total = align_up(L.off_keysB + size_t(n) * key_size);
    return L;
}

// The onesweep sort with 11-bit digits (keys only; 1 < n < 2^30).
template <typename K, int FLOATK>
void run_sort11(hpcs_ctx *c, K *d_keys, int64_t n, hipStream_t st)
{
    constexpr int P = d11_passes<K>();
    const Layout11 L = make_layout11(n, sizeof(K));
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, L.total));
    uint32_t *hist = reinterpret_cast<uint32_t *>(base);
    DevPlan11 *plan = reinterpret_cast<DevPlan11 *>(base + L.off_plan);
    uint32_t *lb[2] = {reinterpret_cast<uint32_t *>(base + L.off_lb0), reinterpret_cast<uint32_t *>(base + L.off_lb1)};
    K *keysB = reinterpret_cast<K *>(base + L.off_keysB);
    launch_fill(c, base, L.zero_bytes, st);
    timed(c, st, "histogram", [&] {
        hipLaunchKernelGGL((k_hist11<K, FLOATK>), dim3(grid_for(uint64_t(n), 256, c->num_cus * 4)), dim3(256), 0, st,
                           d_keys, uint64_t(n), hist);
    });
    timed(c, st, "plan", [&] {
        hipLaunchKernelGGL((k_plan11<K>), dim3(1), dim3(1024), 0, st, hist, plan, uint64_t(n), FLOATK ? 1 : 0);
    });
    for (int p = 0; p < P; ++p) {
        uint32_t *next = (p + 1 < P) ? lb[(p + 1) & 1] : nullptr;
        timed(c, st, "onesweep", [&] {
            hipLaunchKernelGGL((k_onesweep11<K, FLOATK>), dim3(unsigned(L.tiles)), dim3(D11_BLOCK), 0, st, d_keys,
                               keysB, uint32_t(n), p, plan, lb[p & 1], next, c->fallback_rounds);
        });
    }
    timed(c, st, "copyback", [&] {
        hipLaunchKernelGGL((k_copyback11<K>), dim3(grid_for(uint64_t(n), 256, c->num_cus * 8)), dim3(256), 0, st,
                           d_keys, keysB, uint64_t(n), plan);
    });
}

// Bucket tile for the MSD path: about 2.5x the mean bucket (n / 256) where a
// workgroup's LDS holds it; larger buckets take the streaming passes.
template <typename K, int FLOATK>
void launch_sort_buckets(int64_t n, hipStream_t st, K *a, K *b, const DevPlan *plan)
{
    const int64_t want = n / RADIX * 5 / 2;
    if constexpr (sizeof(K) == 4) {
        if (want <= 4096)
            hipLaunchKernelGGL((k_sort_buckets<K, FLOATK, 256, 16>), dim3(RADIX), dim3(256), 0, st, a, b, uint32_t(n),
                               plan);
        else if (want <= 8192)
            hipLaunchKernelGGL((k_sort_buckets<K, FLOATK, 512, 16>), dim3(RADIX), dim3(512), 0, st, a, b, uint32_t(n),
                               plan);
        else
            hipLaunchKernelGGL((k_sort_buckets<K, FLOATK, 1024, 16>), dim3(RADIX), dim3(1024), 0, st, a, b,
                               uint32_t(n), plan);
    } else {
        if (want <= 2048)
            hipLaunchKernelGGL((k_sort_buckets<K, FLOATK, 256, 8>), dim3(RADIX), dim3(256), 0, st, a, b, uint32_t(n),
                               plan);
        else if (want <= 4096)
            hipLaunchKernelGGL((k_sort_buckets<K, FLOATK, 256, 16>), dim3(RADIX), dim3(256), 0, st, a, b, uint32_t(n),
                               plan);
        else
            hipLaunchKernelGGL((k_sort_buckets<K, FLOATK, 512, 16>), dim3(RADIX), dim3(512), 0, st, a, b, uint32_t(n),
                               plan);
    }
}

// One MSD split by the top eight varying bits (onesweep pass -1), then one
// workgroup per bucket (keys without values, n <= MSD_MAX4 / MSD_MAX8).
template <typename K, int FLOATK>
void run_sort_msd(hpcs_ctx *c, K *d_keys, int64_t n, hipStream_t st, const SortParams &prm)
{
    const int shape = prm.shape;
    const int tile_keys = TILE_SHAPES[shape].block * TILE_SHAPES[shape].items;
    const Layout L = make_layout(n, sizeof(K), false, tile_keys, false);
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, L.total));
    DevPlan *plan = reinterpret_cast<DevPlan *>(base + L.off_plan);
    uint32_t *lb0 = reinterpret_cast<uint32_t *>(base + L.off_lb0);
    K *keysB = reinterpret_cast<K *>(base + L.off_keysB);
    K *orslots = reinterpret_cast<K *>(base + L.off_hist); // the histogram area holds the OR slots
    launch_fill(c, base, L.zero_bytes, st);                  // OR slots, plan, look-back words
    const int grid = grid_for(uint64_t(n), MSD_BLOCK * 8, c->num_cus * 4);
    timed(c, st, "msd_or_and", [&] {
        hipLaunchKernelGGL((k_or_and<K, FLOATK>), dim3(grid), dim3(MSD_BLOCK), 0, st, d_keys, uint32_t(n), orslots);
    });
    timed(c, st, "msd_hist", [&] {
        hipLaunchKernelGGL((k_msd_hist<K, FLOATK>), dim3(grid), dim3(MSD_BLOCK), 0, st, d_keys, uint32_t(n), orslots,
                           plan);
    });
    timed(c, st, "plan", [&] {
        hipLaunchKernelGGL((k_plan_msd<K>), dim3(1), dim3(RADIX), 0, st, orslots, plan, uint32_t(n), FLOATK ? 1 : 0);
    });
    timed(c, st, "msd_split", [&] {
        launch_onesweep<K, false, FLOATK>(prm.rank, shape, c->lookback_window, c->fallback_rounds,
                                          dim3(unsigned(L.tiles)), st, d_keys, keysB, nullptr, nullptr, uint32_t(n),
                                          -1, plan, lb0, nullptr, nullptr);
    });
    timed(c, st, "msd_buckets", [&] { launch_sort_buckets<K, FLOATK>(n, st, d_keys, keysB, plan); });
}

void launch_seg_counts(int shape, dim3 grid, hipStream_t st, const uint32_t *a, const uint32_t *b, int pass,
                       const SegPlan *sp, const uint4 *tinfo, uint16_t *counts, const uint32_t *dshift,
                       bool stride = false);
void launch_seg_scatter(int shape, int rank, dim3 grid, hipStream_t st, uint32_t *a, uint32_t *b, int pass,
                        const SegPlan *sp, const uint4 *tinfo, const uint32_t *offsets, const uint32_t *rbase,
                        const uint32_t *sbase, const uint32_t *dshift, bool stride = false);

// Workgroups per CU of the overflow stages' segmented passes (grid-stride
// over their tiles): few enough that the usual empty launch costs ~2 us
// instead of ~14 (one workgroup per tile bound), many enough that a full one
// (keys crowded into few buckets) runs as fast as one workgroup per tile
// (2^28 keys, 90% in one top byte: 3.26 ms against 3.25; 8 per CU: 3.80).
constexpr int OV_GRID_PER_CU = 128;

// Scratch of the MSD-first form (hpcs_hybrid.hpp): the reduce-then-scan
// layout, then the top-byte tile counts, the bucket bounds, level 2's and the
// overflow's segment plans and lists, and the tile records, counts and
// offsets both share (level 2 is done when the overflow starts).
struct HybLayout {
    Layout L;
    size_t off_cnt3, off_begin1, off_sp, off_sp2, off_tile0, off_small, off_medium, off_large, off_tinfo, off_scnt,
        off_soff, off_rbase, off_total, off_sbase, off_dstart, off_pairs, off_medium2, off_large2, off_tile02, total;
    uint64_t max_tiles1, max_tiles2; // tile bounds of level 2 and of the overflow passes
    uint64_t max_ranges, max_large2; // ranges of SEGX_RANGE tile records; overflow segments of more than SEG_MEDIUM keys
    int seg_shape, seg_tile;
};

HybLayout make_hyb_layout(int64_t n, int sort_shape)
{
    HybLayout H{};
    H.L = make_layout(n, 4, false, shape_tile_keys(sort_shape), false);
    H.seg_shape = seg_shape_ok(sort_shape) ? sort_shape : 0;
    H.seg_tile = shape_tile_keys(H.seg_shape);
    const uint64_t tiles = uint64_t(n) / uint64_t(H.seg_tile);
    H.max_tiles1 = tiles + RADIX + 1;                           // the 256 buckets' tiles
    H.max_tiles2 = tiles + uint64_t(n) / uint64_t(SEG_MEDIUM) + 2; // overflow segments exceed SEG_MEDIUM keys
    const uint64_t mt = std::max(H.max_tiles1, H.max_tiles2);
    H.off_cnt3 = H.L.total;
    H.off_begin1 = align_up(H.off_cnt3 + size_t(H.L.tiles) * RADIX * sizeof(uint16_t));
    H.off_sp = align_up(H.off_begin1 + (RADIX + 1) * sizeof(int64_t));
    H.off_sp2 = align_up(H.off_sp + sizeof(SegPlan));
    H.off_tile0 = align_up(H.off_sp2 + sizeof(SegPlan));
    H.off_small = align_up(H.off_tile0 + RADIX * 4);
    H.off_medium = align_up(H.off_small + RADIX * 4);
    H.off_large = align_up(H.off_medium + RADIX * 4);
    H.off_tinfo = align_up(H.off_large + RADIX * 4);
    H.off_scnt = align_up(H.off_tinfo + mt * sizeof(uint4));
    H.off_soff = align_up(H.off_scnt + mt * RADIX * sizeof(uint16_t));
    H.max_ranges = (mt + SEGX_RANGE - 1) / SEGX_RANGE;
    H.max_large2 = uint64_t(n) / uint64_t(SEG_MEDIUM) + 1;
    H.off_rbase = align_up(H.off_soff + mt * RADIX * sizeof(uint32_t));
    H.off_total = align_up(H.off_rbase + H.max_ranges * RADIX * sizeof(uint32_t));
    H.off_sbase = align_up(H.off_total + RADIX * sizeof(uint32_t));
    H.off_dstart = align_up(H.off_sbase + std::max<uint64_t>(RADIX, H.max_large2) * RADIX * sizeof(uint32_t));
    H.off_pairs = align_up(H.off_dstart + size_t(RADIX) * RADIX * sizeof(uint32_t));
    H.off_medium2 = align_up(H.off_pairs + size_t(HYB_SUB) * 2 * sizeof(int64_t));
    H.off_large2 = align_up(H.off_medium2 + size_t(HYB_SUB) * 4);
    H.off_tile02 = align_up(H.off_large2 + size_t(HYB_SUB) * 4);
    H.total = align_up(H.off_tile02 + size_t(HYB_SUB) * 2 * 4);
    return H;
}

// The MSD-first form (pass algorithm 6) for 4-byte unsigned keys without
// values; inputs whose top two bytes do not both vary take the plain
// reduce-then-scan passes inside the same launch sequence (device plan).
void run_sort_hybrid(hpcs_ctx *c, uint32_t *d_keys, int64_t n, hipStream_t st, const SortParams &prm)
{
    using K = uint32_t;
    const int shape = prm.shape, rank_mode = prm.rank;
    const HybLayout H = make_hyb_layout(n, shape);
    const Layout &L = H.L;
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, H.total));
    DevPlan *plan = reinterpret_cast<DevPlan *>(base + L.off_plan);
    uint32_t *offs = reinterpret_cast<uint32_t *>(base + L.off_lb0), *totals = reinterpret_cast<uint32_t *>(base + L.off_lb1);
    uint32_t *dtot = reinterpret_cast<uint32_t *>(base + L.off_dtot);
    uint16_t *counts = reinterpret_cast<uint16_t *>(base + L.off_cnt);
    uint16_t *counts3 = reinterpret_cast<uint16_t *>(base + H.off_cnt3);
    K *keysB = reinterpret_cast<K *>(base + L.off_keysB);
    K *orslots = reinterpret_cast<K *>(base + L.off_hist);
    int64_t *begin1 = reinterpret_cast<int64_t *>(base + H.off_begin1);
    SegPlan *sp = reinterpret_cast<SegPlan *>(base + H.off_sp);
    SegPlan *sp2 = reinterpret_cast<SegPlan *>(base + H.off_sp2);
    uint32_t *tile0 = reinterpret_cast<uint32_t *>(base + H.off_tile0);
    uint32_t *small = reinterpret_cast<uint32_t *>(base + H.off_small);
    uint32_t *medium = reinterpret_cast<uint32_t *>(base + H.off_medium);
    uint32_t *large = reinterpret_cast<uint32_t *>(base + H.off_large);
    uint4 *tinfo = reinterpret_cast<uint4 *>(base + H.off_tinfo);
    uint16_t *scnt = reinterpret_cast<uint16_t *>(base + H.off_scnt);
    uint32_t *soff = reinterpret_cast<uint32_t *>(base + H.off_soff);
    uint32_t *rbase = reinterpret_cast<uint32_t *>(base + H.off_rbase);
    uint32_t *stotal = reinterpret_cast<uint32_t *>(base + H.off_total);
    uint32_t *sbase = reinterpret_cast<uint32_t *>(base + H.off_sbase);
    uint32_t *dstart = reinterpret_cast<uint32_t *>(base + H.off_dstart);
    int64_t *pairs = reinterpret_cast<int64_t *>(base + H.off_pairs);
    uint32_t *medium2 = reinterpret_cast<uint32_t *>(base + H.off_medium2);
    uint32_t *large2 = reinterpret_cast<uint32_t *>(base + H.off_large2);
    uint32_t *tile02 = reinterpret_cast<uint32_t *>(base + H.off_tile02);
    const uint32_t tiles = uint32_t(L.tiles);
    const uint32_t nranges = (tiles + SCAN_RANGE - 1) / SCAN_RANGE;
    const int stile = H.seg_tile;

    launch_fill(c, base, L.off_lb0, st); // OR slots, plan, per-pass digit totals
    timed(c, st, "upsweep", [&] {
        with_shape<4, false>(shape, [&](auto sc) {
            constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
            hipLaunchKernelGGL((k_upsweep<K, 0, T.block, T.items, true>), dim3(tiles), dim3(T.block), 0, st, d_keys,
                               uint32_t(n), counts, orslots, int(aligned16(d_keys)), counts3);
        });
    });
    timed(c, st, "plan", [&] {
        hipLaunchKernelGGL((k_plan_rts<K>), dim3(1), dim3(64), 0, st, orslots, plan, uint64_t(n), 0, 1);
    });
    // the global pass slot (plan pass 3): the first split, or the LSD pass
    // over the lowest varying byte
    {
        constexpr int p = 3;
        timed(c, st, "tile_counts", [&] {
            launch_tile_counts<K, false, 0>(shape, dim3(tiles), st, d_keys, keysB, uint32_t(n), p, plan, counts,
                                            nullptr);
        });
        timed(c, st, "scan", [&] {
            uint32_t *dtot_p = dtot + size_t(p) * RADIX * LINE_WORDS;
            hipLaunchKernelGGL(k_range_totals, dim3(nranges), dim3(RADIX), 0, st, counts, counts3, offs, tiles,
                               totals, dtot_p, p, plan);
            hipLaunchKernelGGL(k_scan_columns, dim3(RADIX), dim3(1024), 0, st, totals, nranges, dtot_p, p, plan);
        });
        timed(c, st, "scatter", [&] {
            launch_onesweep<K, false, 0>(rank_mode, shape, 0, c->fallback_rounds, dim3(tiles), st, d_keys, keysB,
                                         nullptr, nullptr, uint32_t(n), p, plan, offs, totals, nullptr);
        });
    }
    // level 2: the top-byte buckets in B, split by byte 2 into A (buckets of
    // at most SEG_SMALL keys sorted whole into A)
    timed(c, st, "hyb_begin", [&] {
        hipLaunchKernelGGL(k_hyb_begin1, dim3(1), dim3(RADIX), 0, st, plan, uint64_t(n), begin1, sp2);
    });
    timed(c, st, "seg_plan", [&] {
        hipLaunchKernelGGL(k_seg_plan, dim3(1), dim3(1024), 0, st, begin1, RADIX, n, sp, tile0, small, medium, large,
                           H.max_tiles1, SEG_SMALL, stile, 1);
    });
    timed(c, st, "seg_tiles", [&] {
        hipLaunchKernelGGL(k_seg_tiles, dim3(RADIX), dim3(256), 0, st, begin1, sp, tile0, large, tinfo, stile,
                           int64_t(0));
    });
    timed(c, st, "seg_small", [&] {
        hipLaunchKernelGGL((k_seg_small<SEG_BLOCK>), dim3(RADIX), dim3(SEG_BLOCK), 0, st, keysB, d_keys, begin1, sp,
                           small);
    });
    timed(c, st, "seg_counts", [&] {
        launch_seg_counts(H.seg_shape, dim3(unsigned(H.max_tiles1)), st, keysB, d_keys, 2, sp, tinfo, scnt,
                          &plan->hyb_shift2);
    });
    // the large segments' scan: range sums, column scan, segment bases
    auto segx_scan = [&](const int64_t *bg, const SegPlan *p, const uint32_t *t0, const uint32_t *lst, uint64_t mt,
                         uint64_t nlarge_max, uint32_t *ds) {
        timed(c, st, "seg_scan", [&] {
            hipLaunchKernelGGL(k_segx_ranges, dim3(unsigned((mt + SEGX_RANGE - 1) / SEGX_RANGE)), dim3(RADIX), 0, st,
                               p, scnt, soff, rbase);
            hipLaunchKernelGGL(k_segx_columns, dim3(RADIX), dim3(1024), 0, st, p, rbase, stotal);
            hipLaunchKernelGGL(k_segx_bases, dim3(unsigned(std::min<uint64_t>(nlarge_max, uint64_t(c->num_cus) * 8))),
                               dim3(RADIX), 0, st, bg, p, t0, lst, soff, rbase, stotal, sbase, ds, stile);
        });
    };
    segx_scan(begin1, sp, tile0, large, H.max_tiles1, RADIX, dstart);
    timed(c, st, "seg_scatter", [&] {
        launch_seg_scatter(H.seg_shape, rank_mode, dim3(unsigned(H.max_tiles1)), st, keysB, d_keys, 2, sp, tinfo,
                           soff, rbase, sbase, &plan->hyb_shift2);
    });
    // level 3: every sub-bucket in LDS, in place
    timed(c, st, "hyb_local", [&] {
        hipLaunchKernelGGL((k_hyb_local<HYB_BLOCK, HYB_ITEMS>), dim3(HYB_SUB), dim3(HYB_BLOCK), 0, st, d_keys, begin1,
                           plan, sp, large, dstart, sp2, pairs, medium2, large2, tile02, stile);
    });
    // overflow: medium sub-buckets in LDS, large ones by bytes 0 and 1
    timed(c, st, "hyb_medium", [&] {
        hipLaunchKernelGGL((k_seg_small<SEG_MEDIUM_BLOCK>), dim3(c->num_cus), dim3(SEG_MEDIUM_BLOCK), 0, st, d_keys,
                           d_keys, pairs, sp2, medium2);
    });
    timed(c, st, "seg_tiles", [&] {
        hipLaunchKernelGGL(k_seg_tiles, dim3(c->num_cus * 8), dim3(256), 0, st, pairs, sp2, tile02, large2, tinfo,
                           stile, int64_t(0));
    });
    // the overflow passes walk their tiles grid-stride from OV_GRID_PER_CU
    // workgroups per CU (usually they have no tiles at all)
    const dim3 ov_grid(unsigned(std::min<uint64_t>(H.max_tiles2, uint64_t(c->num_cus) * OV_GRID_PER_CU)));
    for (int p = 0; p < 2; ++p) {
        timed(c, st, "seg_counts", [&] {
            launch_seg_counts(H.seg_shape, ov_grid, st, d_keys, keysB, p, sp2, tinfo, scnt, &plan->hyb_ov_shift[p],
                              true);
        });
        segx_scan(pairs, sp2, tile02, large2, H.max_tiles2, H.max_large2, nullptr);
        timed(c, st, "seg_scatter", [&] {
            launch_seg_scatter(H.seg_shape, rank_mode, ov_grid, st, d_keys, keysB, p, sp2, tinfo, soff, rbase, sbase,
                               &plan->hyb_ov_shift[p], true);
        });
    }
    timed(c, st, "copyback", [&] {
        hipLaunchKernelGGL((k_copyback<K>), dim3(grid_for(uint64_t(n), 256, c->num_cus * 8)), dim3(256), 0, st,
                           d_keys, keysB, nullptr, nullptr, uint64_t(n), plan);
    });
}

// Scratch memory a caller hands to run_sort instead of the context's pool
// (the no-allocation entry point): its device address and size.
struct ExtScratch {
    void *ptr;
    size_t bytes;
};

template <typename K, bool HAS_VALS, int FLOATK>
void run_sort(hpcs_ctx *c, K *d_keys, uint32_t *d_vals, int64_t n, hipStream_t st, const ExtScratch *ext = nullptr)
{
    if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
    if (n > 0 && !d_keys) fail(HPCS_E_INVALID_ARG, "keys is null");
    if (HAS_VALS && n > 0 && !d_vals) fail(HPCS_E_INVALID_ARG, "values is null");
    if (n <= 1) return;
    if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
    if (reinterpret_cast<uintptr_t>(d_keys) % sizeof(K)) fail(HPCS_E_INVALID_ARG, "keys pointer is misaligned");
    constexpr int P = int(sizeof(K));
    SortParams prm = resolve_params(c, int(sizeof(K)), HAS_VALS, n);
    if (ext) {
        // caller-provided scratch: the 8-bit passes (whose layout is
        // make_layout's), never the 11-bit or MSD-split layouts
        prm.bits = 8;
        if (prm.algo == PASS_MSD || prm.algo == PASS_MERGE || prm.algo == PASS_HYBRID)
            prm.algo = n <= (sizeof(K) == 4 ? BLOCK_SORT_MAX4 : BLOCK_SORT_MAX8) ? PASS_BLOCK
                       : n < (sizeof(K) == 4 ? (int64_t(1) << 19) : (int64_t(1) << 21)) ? PASS_ONESWEEP
                                                                                         : PASS_RTS;
    }
    if (prm.algo == PASS_HYBRID) {
        if constexpr (std::is_same<K, uint32_t>::value && FLOATK == 0 && !HAS_VALS)
            return run_sort_hybrid(c, d_keys, n, st, prm);
        prm.algo = PASS_RTS; // float and signed keys: the same passes without the bucket stages
    }
    if (prm.algo == PASS_MERGE) return run_sort_merge<K, HAS_VALS, FLOATK>(c, d_keys, d_vals, n, st);
    if (prm.algo == PASS_BLOCK) {
        static_assert(BLOCK_SORT_MAX4 == 1024 * 16 && BLOCK_SORT_MAX8 == 1024 * 8, "k_sort_block sizes");
        return run_sort_block<K, HAS_VALS, FLOATK>(c, d_keys, d_vals, n, st);
    }
    if constexpr (!HAS_VALS) {
        if (prm.bits == 11) return run_sort11<K, FLOATK>(c, d_keys, n, st);
        if (prm.algo == PASS_MSD) return run_sort_msd<K, FLOATK>(c, d_keys, n, st, prm);
    }
    const int shape = prm.shape;
    const int tile_keys = TILE_SHAPES[shape].block * TILE_SHAPES[shape].items;
    const bool want_dig = digit_bytes_on(c, sizeof(K), HAS_VALS);
    const Layout L = make_layout(n, sizeof(K), HAS_VALS, tile_keys, want_dig);
    if (ext && L.total > ext->bytes) fail(HPCS_E_UNSUPPORTED, "the sort's scratch exceeds the static scratch");
    char *base = ext ? static_cast<char *>(ext->ptr) : static_cast<char *>(ensure(c->scratch, c->scratch_cap, L.total));
    uint32_t *hist = reinterpret_cast<uint32_t *>(base + L.off_hist);
    DevPlan *plan = reinterpret_cast<DevPlan *>(base + L.off_plan);
    uint32_t *lb[2] = {reinterpret_cast<uint32_t *>(base + L.off_lb0), reinterpret_cast<uint32_t *>(base + L.off_lb1)};
    uint32_t *dtot = reinterpret_cast<uint32_t *>(base + L.off_dtot); // per-pass digit totals, zero-filled
    K *keysB = reinterpret_cast<K *>(base + L.off_keysB);
    uint32_t *valsB = HAS_VALS ? reinterpret_cast<uint32_t *>(base + L.off_valsB) : nullptr;

    const uint32_t tiles = uint32_t(L.tiles);
    const uint32_t nranges = (tiles + SCAN_RANGE - 1) / SCAN_RANGE;
    // look-back words carry 30-bit counts: larger sorts take the
    // reduce-then-scan passes whatever the option or table says
    const int algo = (prm.algo == PASS_ONESWEEP && n >= LOOKBACK_MAX_KEYS) ? PASS_RTS : prm.algo;
    const int rank_mode = prm.rank;
    uint8_t *dig = want_dig ? reinterpret_cast<uint8_t *>(base + L.off_dig) : nullptr;
    if (algo == PASS_RTS && tiles <= SCAN_DIGIT_MAX_TILES) {
        // few tiles: launches, not bytes, set the time.  One histogram of
        // every digit gives the plan (digit starts, trivial passes); per
        // pass the tile counts, one scan kernel and the scatter.
        launch_fill(c, base, L.off_lb0, st);
        launch_histogram<K, FLOATK>(c, d_keys, uint64_t(n), hist, st);
        timed(c, st, "plan", [&] {
            hipLaunchKernelGGL((k_plan<P>), dim3(1), dim3(RADIX), 0, st, hist, plan, uint64_t(n), FLOATK ? 1 : 0);
        });
        // 16-bit tile counts, then 32-bit global offsets in look-back region 0
        uint16_t *counts = reinterpret_cast<uint16_t *>(base + L.off_cnt);
        uint32_t *offs = lb[0];
        for (int p = 0; p < P; ++p) {
            timed(c, st, "tile_counts", [&] {
                launch_tile_counts<K, HAS_VALS, FLOATK>(shape, dim3(tiles), st, d_keys, keysB, uint32_t(n), p, plan,
                                                        counts, nullptr);
            });
            timed(c, st, "scan", [&] {
                hipLaunchKernelGGL(k_scan_digit, dim3(RADIX), dim3(1024), 0, st, counts, offs, tiles, p, plan);
            });
            timed(c, st, "scatter", [&] {
                launch_onesweep<K, HAS_VALS, FLOATK>(rank_mode, shape, 0, c->fallback_rounds, dim3(tiles), st, d_keys,
                                                     keysB, d_vals, valsB, uint32_t(n), p, plan, offs, nullptr,
                                                     nullptr);
            });
        }
        timed(c, st, "copyback", [&] {
            hipLaunchKernelGGL((k_copyback<K>), dim3(grid_for(uint64_t(n), 256, c->num_cus * 8)), dim3(256), 0, st,
                               d_keys, keysB, HAS_VALS ? d_vals : nullptr, valsB, uint64_t(n), plan);
        });
        return;
    }
    if (algo == PASS_RTS) {
        // the histogram area holds the OR slots; counts in look-back region 0
        launch_fill(c, base, L.off_lb0, st);
        K *orslots = reinterpret_cast<K *>(hist);
        timed(c, st, "upsweep", [&] {
            launch_upsweep<K, HAS_VALS, FLOATK>(shape, dim3(tiles), st, d_keys, uint32_t(n),
                                                reinterpret_cast<uint16_t *>(base + L.off_cnt), orslots);
        });
        timed(c, st, "plan", [&] {
            hipLaunchKernelGGL((k_plan_rts<K>), dim3(1), dim3(64), 0, st, orslots, plan, uint64_t(n),
                               FLOATK ? 1 : 0);
        });
    } else {
        launch_fill(c, base, L.zero_bytes, st);
        launch_histogram<K, FLOATK>(c, d_keys, uint64_t(n), hist, st);
        timed(c, st, "plan", [&] {
            hipLaunchKernelGGL((k_plan<P>), dim3(1), dim3(RADIX), 0, st, hist, plan, uint64_t(n), FLOATK ? 1 : 0);
        });
    }
    for (int p = 0; p < P; ++p) {
        if (algo == PASS_RTS) {
            // 16-bit tile counts; in-range offsets in look-back region 0,
            // range sums in region 1
            uint16_t *counts = reinterpret_cast<uint16_t *>(base + L.off_cnt);
            uint32_t *offs = lb[0], *totals = lb[1];
            if (p > 0)
                timed(c, st, "tile_counts", [&] {
                    launch_tile_counts<K, HAS_VALS, FLOATK>(shape, dim3(tiles), st, d_keys, keysB, uint32_t(n), p,
                                                            plan, counts, dig);
                });
            timed(c, st, "scan", [&] {
                uint32_t *dtot_p = dtot + size_t(p) * RADIX * LINE_WORDS;
                hipLaunchKernelGGL(k_range_totals, dim3(nranges), dim3(RADIX), 0, st, counts, counts, offs, tiles,
                                   totals, dtot_p, p, plan);
                hipLaunchKernelGGL(k_scan_columns, dim3(RADIX), dim3(1024), 0, st, totals, nranges, dtot_p, p, plan);
            });
            timed(c, st, "scatter", [&] {
                // the scatter adds its range's base to its in-range offsets
                launch_onesweep<K, HAS_VALS, FLOATK>(rank_mode, shape, 0, c->fallback_rounds, dim3(tiles), st, d_keys,
                                                     keysB, d_vals, valsB, uint32_t(n), p, plan, offs, totals, dig);
            });
            continue;
        }
        uint32_t *next = (p + 1 < P) ? lb[(p + 1) & 1] : nullptr;
        timed(c, st, "onesweep", [&] {
            launch_onesweep<K, HAS_VALS, FLOATK>(rank_mode, shape, c->lookback_window, c->fallback_rounds,
                                                 dim3(unsigned(L.tiles)), st, d_keys, keysB, d_vals, valsB,
                                                 uint32_t(n), p, plan, lb[p & 1], next, nullptr);
        });
    }
    timed(c, st, "copyback", [&] {
        hipLaunchKernelGGL((k_copyback<K>), dim3(grid_for(uint64_t(n), 256, c->num_cus * 8)), dim3(256), 0, st,
                           d_keys, keysB, HAS_VALS ? d_vals : nullptr, valsB, uint64_t(n), plan);
    });
}

// Scratch bytes run_sort takes for n keys with the context's options and
// table (ext: the plan a caller-provided scratch gets: 8-bit passes, no MSD
// split); the one-workgroup sort takes none.
template <typename K, bool HAS_VALS>
size_t sort_scratch_bytes(const hpcs_ctx *c, int64_t n, bool ext)
{
    if (n <= 1) return 0;
    SortParams prm = resolve_params(c, int(sizeof(K)), HAS_VALS, n);
    if (ext) {
        prm.bits = 8;
        if (prm.algo == PASS_MSD || prm.algo == PASS_MERGE || prm.algo == PASS_HYBRID)
            prm.algo = n <= (sizeof(K) == 4 ? BLOCK_SORT_MAX4 : BLOCK_SORT_MAX8) ? PASS_BLOCK : PASS_RTS;
    }
    if (prm.algo == PASS_BLOCK) return 0;
    if (prm.algo == PASS_HYBRID && sizeof(K) == 4 && !HAS_VALS) return make_hyb_layout(n, prm.shape).total;
    if (prm.algo == PASS_MERGE) return merge_sort_scratch_bytes(n, sizeof(K), HAS_VALS);
    if (!HAS_VALS && prm.bits == 11) return make_layout11(n, sizeof(K)).total;
    return make_layout(n, sizeof(K), HAS_VALS, shape_tile_keys(prm.shape),
                       prm.algo != PASS_MSD && digit_bytes_on(c, sizeof(K), HAS_VALS))
        .total;
}

// The memory-tight keys-only sort (hpcs_merge.hpp): each half sorted in place
// by the LSD passes, then the halves merged with half the keys of scratch.
template <typename K, int FLOATK>
void run_sort_tight(hpcs_ctx *c, K *d_keys, int64_t n, hipStream_t st)
{
    const int64_t h = (n + 1) / 2, nb = n - h;
    const size_t merge_bytes = align_up(size_t(h) * sizeof(K)) + align_up(sizeof(MergeSplit));
    // halves whose tuned sort is the block-sort + merge form take it (its
    // scratch is one half's keys); others take the 8-bit passes' layout
    const bool merge_halves = resolve_params(c, int(sizeof(K)), false, h).algo == PASS_MERGE &&
                              resolve_params(c, int(sizeof(K)), false, nb).algo == PASS_MERGE;
    const size_t halves = merge_halves ? merge_sort_scratch_bytes(h, sizeof(K), false)
                                       : std::max(sort_scratch_bytes<K, false>(c, h, true),
                                                  sort_scratch_bytes<K, false>(c, nb, true));
    const size_t need = std::max(merge_bytes, halves);
    if (c->scratch_limit && need > c->scratch_limit)
        fail(HPCS_E_BUDGET, "the sort needs " + std::to_string(need) + " bytes of scratch even with halves merged (" +
                                "limit " + std::to_string(c->scratch_limit) + ")");
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, need));
    if (merge_halves) {
        run_sort_merge<K, false, FLOATK>(c, d_keys, nullptr, h, st);
        run_sort_merge<K, false, FLOATK>(c, d_keys + h, nullptr, nb, st);
    } else {
        const ExtScratch ext{base, std::max(need, c->scratch_cap)};
        run_sort<K, false, FLOATK>(c, d_keys, nullptr, h, st, &ext);
        run_sort<K, false, FLOATK>(c, d_keys + h, nullptr, nb, st, &ext);
    }
    K *S = reinterpret_cast<K *>(base);
    MergeSplit *sp = reinterpret_cast<MergeSplit *>(base + align_up(size_t(h) * sizeof(K)));
    K *R = d_keys + h;
    timed(c, st, "merge", [&] {
        hipLaunchKernelGGL((k_copy_keys<K>), dim3(grid_for(uint64_t(h), 256, c->num_cus * 8)), dim3(256), 0, st, S,
                           d_keys, uint64_t(h));
    });
    const dim3 grid(unsigned((h + MERGE_TILE - 1) / MERGE_TILE));
    timed(c, st, "merge", [&] {
        hipLaunchKernelGGL((k_merge_split<K, FLOATK>), dim3(1), dim3(64), 0, st, S, uint64_t(h), R, uint64_t(nb), sp);
    });
    timed(c, st, "merge", [&] {
        hipLaunchKernelGGL((k_merge<K, FLOATK>), grid, dim3(MERGE_BLOCK), 0, st, S, uint64_t(h), R, uint64_t(nb), sp, 0,
                           d_keys);
    });
    timed(c, st, "merge", [&] {
        hipLaunchKernelGGL((k_merge_tail<K>), dim3(grid_for(uint64_t(nb), 256, c->num_cus * 8)), dim3(256), 0, st, R,
                           uint64_t(nb), sp, S);
    });
    timed(c, st, "merge", [&] {
        hipLaunchKernelGGL((k_merge<K, FLOATK>), grid, dim3(MERGE_BLOCK), 0, st, S, uint64_t(h), R, uint64_t(nb), sp, 1,
                           R);
    });
}

// A sort with the context's scratch limit: keys-only sorts whose full
// scratch exceeds the limit, or whose scratch allocation fails, take the
// memory-tight form (as SortRadixInPlaceAdaptive falls back to an in-place
// sort on OutOfMemoryException, RadixSortMsd.cs:725-770); sorts with values
// beyond the limit fail with HPCS_E_BUDGET before any work; others run
// directly.
template <typename K, bool HAS_VALS, int FLOATK>
void run_sort_limited(hpcs_ctx *c, K *d_keys, uint32_t *d_vals, int64_t n, hipStream_t st)
{
    if constexpr (!HAS_VALS) {
        if (n > 1 && n < MAX_KEYS && d_keys && reinterpret_cast<uintptr_t>(d_keys) % sizeof(K) == 0) {
            if (c->scratch_limit && sort_scratch_bytes<K, false>(c, n, false) > c->scratch_limit)
                return run_sort_tight<K, FLOATK>(c, d_keys, n, st);
            try {
                return run_sort<K, false, FLOATK>(c, d_keys, nullptr, n, st);
            } catch (const Err &e) {
                if (e.code != HPCS_E_OOM) throw;
                (void)hipGetLastError();
            }
            return run_sort_tight<K, FLOATK>(c, d_keys, n, st);
        }
    } else {
        if (c->scratch_limit && n > 1 && n < MAX_KEYS) {
            const size_t need = sort_scratch_bytes<K, true>(c, n, false);
            if (need > c->scratch_limit)
                fail(HPCS_E_BUDGET, "the sort of keys with values needs " + std::to_string(need) +
                                        " bytes of scratch (limit " + std::to_string(c->scratch_limit) + ")");
        }
    }
    run_sort<K, HAS_VALS, FLOATK>(c, d_keys, d_vals, n, st);
}

// Batched segments: scratch [SegPlan | seg_tile0 | small | medium | large |
// tile records | 16-bit counts | offsets | keys B].  The plan's word `bad`
// counts rejected segments.
//
// With the MSD-first form (segments longer than msd_min > 0 keys) its
// bucket bounds and overflow follow: [overflow plan | digit positions |
// bucket bounds | pairs | medium list | large list | first tiles]; at most
// n / (msd_min + 1) segments take the form (256 buckets each), and an
// overflow bucket holds more than HYB_MAX_BIN keys, so at most
// n / (HYB_MAX_BIN + 1) of them, each large one more than SEG_MEDIUM.
struct SegLayout {
    size_t off_plan, off_tile0, off_small, off_medium, off_large, off_tinfo, off_counts, off_offsets, off_keysB,
        off_sp2, off_shifts, off_bounds, off_pairs, off_medium2, off_large2, off_tile02, total;
    uint64_t max_tiles;          // tile records: the large segments' or the overflow's, the larger
    uint64_t max_pairs, max_large2, max_msd;
};

SegLayout seg_layout(int64_t n, int nseg, int tile, int64_t msd_min)
{
    const bool msd = msd_min > 0;
    SegLayout L{};
    L.max_tiles = uint64_t(n) / uint64_t(tile) + uint64_t(nseg) + 1; // tiles of the large segments, at most
    L.max_pairs = uint64_t(n) / uint64_t(HYB_MAX_BIN + 1) + 1;
    L.max_large2 = uint64_t(n) / uint64_t(SEG_MEDIUM + 1) + 1;
    if (msd) L.max_tiles = std::max(L.max_tiles, uint64_t(n) / uint64_t(tile) + L.max_large2 + 1);
    L.off_plan = 0;
    L.off_tile0 = align_up(sizeof(SegPlan));
    L.off_small = align_up(L.off_tile0 + size_t(nseg) * 4);
    L.off_medium = align_up(L.off_small + size_t(nseg) * 4);
    L.off_large = align_up(L.off_medium + size_t(nseg) * 4);
    L.off_tinfo = align_up(L.off_large + size_t(nseg) * 4);
    L.off_counts = align_up(L.off_tinfo + L.max_tiles * sizeof(uint4));
    L.off_offsets = align_up(L.off_counts + L.max_tiles * RADIX * sizeof(uint16_t));
    L.off_keysB = align_up(L.off_offsets + L.max_tiles * RADIX * sizeof(uint32_t));
    L.total = align_up(L.off_keysB + size_t(n) * 4);
    if (msd) {
        L.off_sp2 = L.total;
        L.off_shifts = align_up(L.off_sp2 + sizeof(SegPlan));
        L.max_msd = uint64_t(n) / uint64_t(msd_min + 1) + 1;
        L.off_bounds = align_up(L.off_shifts + 4 * sizeof(uint32_t));
        L.off_pairs = align_up(L.off_bounds + L.max_msd * RADIX * sizeof(uint2));
        L.off_medium2 = align_up(L.off_pairs + L.max_pairs * 2 * sizeof(int64_t));
        L.off_large2 = align_up(L.off_medium2 + L.max_pairs * 4);
        L.off_tile02 = align_up(L.off_large2 + L.max_large2 * 4);
        L.total = align_up(L.off_tile02 + L.max_pairs * 2 * 4);
    }
    return L;
}

// The batched sort's knobs for n keys in all: options set explicitly win,
// then the tuning table's `segs` row for the smallest size class that holds
// n (the largest class beyond it), then the defaults.
struct SegParams {
    int small_max, rank, shape;
};

SegParams resolve_seg_params(const hpcs_ctx *c, int64_t n)
{
    SegParams p{c->seg_small, c->seg_rank, c->seg_shape};
    const TuneEntry *best = nullptr, *largest = nullptr;
    for (const TuneEntry &e : c->tune) {
        if (e.key_bytes != 0) continue;
        if (!largest || e.max_n > largest->max_n) largest = &e;
        if (e.max_n >= n && (!best || e.max_n < best->max_n)) best = &e;
    }
    if (!best) best = largest;
    if (best) {
        if (!c->set_seg_small) p.small_max = best->seg_small;
        if (!c->set_seg_rank) p.rank = best->rank;
        if (!c->set_seg_shape && seg_shape_ok(best->shape)) p.shape = best->shape;
    }
    return p;
}

// Runs f with the large segments' tile shape as a compile-time row of
// TILE_SHAPES (the rows seg_shape_ok admits); any other index fails.
template <int S = 0, typename F>
void with_seg_shape(int shape, F &&f)
{
    if constexpr (S < NUM_TILE_SHAPES) {
        if (shape == S) {
            if constexpr (seg_shape_ok(S)) {
                f(std::integral_constant<int, S>{});
                return;
            } else {
                fail(HPCS_E_UNSUPPORTED, "tile shape " + std::to_string(shape) + " does not serve the segments");
            }
        }
        with_seg_shape<S + 1>(shape, std::forward<F>(f));
    } else {
        fail(HPCS_E_UNSUPPORTED, "unknown tile shape " + std::to_string(shape));
    }
}

// Segmented passes, counting and ranking skewed waves by peeling (runs of
// equal digits: crowded buckets, keys that share their top byte; no cost on
// spread keys).  `stride`: the grid-stride instantiation, for grids smaller
// than the tile count (the overflow stages).
void launch_seg_counts(int shape, dim3 grid, hipStream_t st, const uint32_t *a, const uint32_t *b, int pass,
                       const SegPlan *sp, const uint4 *tinfo, uint16_t *counts, const uint32_t *dshift, bool stride)
{
    with_seg_shape(shape, [&](auto sc) {
        constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
        if (stride)
            hipLaunchKernelGGL((k_seg_counts<T.block, T.items, true, true>), grid, dim3(T.block), 0, st, a, b, pass,
                               sp, tinfo, counts, dshift);
        else
            hipLaunchKernelGGL((k_seg_counts<T.block, T.items, true>), grid, dim3(T.block), 0, st, a, b, pass, sp,
                               tinfo, counts, dshift);
    });
}

template <bool STRIDE>
void launch_seg_scatter_t(int shape, int rank, dim3 grid, hipStream_t st, uint32_t *a, uint32_t *b, int pass,
                          const SegPlan *sp, const uint4 *tinfo, const uint32_t *offsets, const uint32_t *rbase,
                          const uint32_t *sbase, const uint32_t *dshift)
{
    with_seg_shape(shape, [&](auto sc) {
        constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
        if (rank == RANK_BALLOT)
            hipLaunchKernelGGL((k_seg_scatter<T.block, T.items, RANK_BALLOT, true, STRIDE>), grid, dim3(T.block), 0, st,
                               a, b, pass, sp, tinfo, offsets, rbase, sbase, dshift);
        else if (rank == RANK_MIX2)
            hipLaunchKernelGGL((k_seg_scatter<T.block, T.items, RANK_MIX2, true, STRIDE>), grid, dim3(T.block), 0, st, a,
                               b, pass, sp, tinfo, offsets, rbase, sbase, dshift);
        else if (rank == RANK_MIX4)
            hipLaunchKernelGGL((k_seg_scatter<T.block, T.items, RANK_MIX4, true, STRIDE>), grid, dim3(T.block), 0, st, a,
                               b, pass, sp, tinfo, offsets, rbase, sbase, dshift);
        else if (rank == RANK_LDSOR)
            hipLaunchKernelGGL((k_seg_scatter<T.block, T.items, RANK_LDSOR, true, STRIDE>), grid, dim3(T.block), 0, st,
                               a, b, pass, sp, tinfo, offsets, rbase, sbase, dshift);
        else
            fail(HPCS_E_UNSUPPORTED, "unknown ranking mode " + std::to_string(rank));
    });
}

// `stride`: as for launch_seg_counts.
void launch_seg_scatter(int shape, int rank, dim3 grid, hipStream_t st, uint32_t *a, uint32_t *b, int pass,
                        const SegPlan *sp, const uint4 *tinfo, const uint32_t *offsets, const uint32_t *rbase,
                        const uint32_t *sbase, const uint32_t *dshift, bool stride)
{
    if (stride) return launch_seg_scatter_t<true>(shape, rank, grid, st, a, b, pass, sp, tinfo, offsets, rbase, sbase, dshift);
    launch_seg_scatter_t<false>(shape, rank, grid, st, a, b, pass, sp, tinfo, offsets, rbase, sbase, dshift);
}

void run_segments(hpcs_ctx *c, uint32_t *d_keys, int64_t n, const int64_t *d_begin, int nseg, hipStream_t st)
{
    if (n < 0 || nseg < 0) fail(HPCS_E_INVALID_ARG, "n and nseg must be non-negative");
    if (nseg == 0 || n <= 1) return;
    if (!d_keys || !d_begin) fail(HPCS_E_INVALID_ARG, "keys or segment offsets are null");
    if (n >= (int64_t(1) << 32)) fail(HPCS_E_UNSUPPORTED, "segments must lie below 2^32 keys");
    const SegParams prm = resolve_seg_params(c, n);
    const int tile = shape_tile_keys(prm.shape);
    const int64_t msd_min = c->seg_msd; // segments longer than this: the MSD-first form
    const bool msd = msd_min > 0 && n > msd_min;
    const SegLayout L = seg_layout(n, nseg, tile, msd ? msd_min : int64_t(0));
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, L.total));
    SegPlan *sp = reinterpret_cast<SegPlan *>(base + L.off_plan);
    uint32_t *tile0 = reinterpret_cast<uint32_t *>(base + L.off_tile0);
    uint32_t *small = reinterpret_cast<uint32_t *>(base + L.off_small);
    uint32_t *medium = reinterpret_cast<uint32_t *>(base + L.off_medium);
    uint32_t *large = reinterpret_cast<uint32_t *>(base + L.off_large);
    uint4 *tinfo = reinterpret_cast<uint4 *>(base + L.off_tinfo);
    uint16_t *counts = reinterpret_cast<uint16_t *>(base + L.off_counts);
    uint32_t *offsets = reinterpret_cast<uint32_t *>(base + L.off_offsets);
    uint32_t *keysB = reinterpret_cast<uint32_t *>(base + L.off_keysB);
    timed(c, st, "seg_plan", [&] {
        hipLaunchKernelGGL(k_seg_plan, dim3(1), dim3(1024), 0, st, d_begin, nseg, n, sp, tile0, small, medium,
                           large, L.max_tiles, prm.small_max, tile, 2);
    });
    const int list_grid = std::min(nseg, c->num_cus * 8);
    timed(c, st, "seg_tiles", [&] {
        hipLaunchKernelGGL(k_seg_tiles, dim3(list_grid), dim3(256), 0, st, d_begin, sp, tile0, large, tinfo, tile,
                           msd ? msd_min : int64_t(0));
    });
    timed(c, st, "seg_small", [&] {
        hipLaunchKernelGGL((k_seg_small<SEG_BLOCK>), dim3(list_grid), dim3(SEG_BLOCK), 0, st, d_keys, d_keys,
                           d_begin, sp, small);
    });
    // the medium list is empty when small_max is SEG_SMALL (its workgroups
    // return at once; the launch sequence stays the same)
    timed(c, st, "seg_medium", [&] {
        hipLaunchKernelGGL((k_seg_small<SEG_MEDIUM_BLOCK>), dim3(std::min(nseg, c->num_cus)), dim3(SEG_MEDIUM_BLOCK),
                           0, st, d_keys, d_keys, d_begin, sp, medium);
    });
    // one workgroup per tile of the large segments (the device knows how
    // many; the grid is the bound and spare workgroups return)
    const dim3 tile_grid(unsigned(L.max_tiles));
    for (int p = 0; p < 4; ++p) {
        timed(c, st, "seg_counts", [&] {
            launch_seg_counts(prm.shape, tile_grid, st, d_keys, keysB, p, sp, tinfo, counts, nullptr);
        });
        // the MSD-first segments keep pass 0's offsets (their bucket bounds)
        const int64_t skip = (msd && p > 0) ? msd_min : INT64_MAX;
        timed(c, st, "seg_scan", [&] {
            hipLaunchKernelGGL(k_seg_scan, dim3(list_grid), dim3(SEG_SCAN_GROUPS * SEG_SCAN_PAIRS), 0, st, d_begin, sp,
                               tile0, large, counts, offsets, tile, skip);
        });
        timed(c, st, "seg_scatter", [&] {
            launch_seg_scatter(prm.shape, prm.rank, tile_grid, st, d_keys, keysB, p, sp, tinfo, offsets, nullptr,
                               nullptr, nullptr);
        });
    }
    if (!msd) return;
    // the MSD-first segments: 256 top-byte buckets each in B, sorted into A
    // (hpcs_hybrid.hpp); overflow buckets by one 1024-thread workgroup in LDS
    // or by the LSD passes over bytes 0, 1, 2 (B -> A -> B -> A)
    SegPlan *sp2 = reinterpret_cast<SegPlan *>(base + L.off_sp2);
    uint32_t *shifts = reinterpret_cast<uint32_t *>(base + L.off_shifts);
    int64_t *pairs = reinterpret_cast<int64_t *>(base + L.off_pairs);
    uint32_t *medium2 = reinterpret_cast<uint32_t *>(base + L.off_medium2);
    uint32_t *large2 = reinterpret_cast<uint32_t *>(base + L.off_large2);
    uint32_t *tile02 = reinterpret_cast<uint32_t *>(base + L.off_tile02);
    uint2 *bounds = reinterpret_cast<uint2 *>(base + L.off_bounds);
    timed(c, st, "seg_msd_init", [&] {
        hipLaunchKernelGGL(k_segmsd_init, dim3(1), dim3(WAVE), 0, st, sp2, shifts);
        hipLaunchKernelGGL(k_segmsd_bounds, dim3(list_grid), dim3(RADIX), 0, st, d_begin, sp, tile0, large, offsets,
                           msd_min, sp2, bounds);
    });
    // one resident round of workgroups (grid-stride over the buckets)
    static const int occ = [] { // resident bucket workgroups per CU, initialised once
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_segmsd_local<SEGMSD_BLOCK, SEGMSD_ITEMS>, SEGMSD_BLOCK, 0) !=
                hipSuccess ||
            o < 1)
            o = 1;
        return o;
    }();
    timed(c, st, "seg_msd_local", [&] {
        hipLaunchKernelGGL((k_segmsd_local<SEGMSD_BLOCK, SEGMSD_ITEMS>), dim3(c->num_cus * occ), dim3(SEGMSD_BLOCK), 0, st, keysB,
                           d_keys, bounds, sp2, pairs, medium2, large2, tile02, tile);
    });
    timed(c, st, "seg_medium", [&] {
        hipLaunchKernelGGL((k_seg_small<SEG_MEDIUM_BLOCK>), dim3(c->num_cus), dim3(SEG_MEDIUM_BLOCK), 0, st, keysB,
                           d_keys, pairs, sp2, medium2);
    });
    timed(c, st, "seg_tiles", [&] {
        hipLaunchKernelGGL(k_seg_tiles, dim3(c->num_cus * 8), dim3(256), 0, st, pairs, sp2, tile02, large2, tinfo,
                           tile, int64_t(0));
    });
    const int list_grid2 = int(std::min<uint64_t>(L.max_large2, uint64_t(c->num_cus) * 8));
    // grid-stride over the overflow tiles (usually none)
    const dim3 ov_grid(unsigned(std::min<uint64_t>(L.max_tiles, uint64_t(c->num_cus) * OV_GRID_PER_CU)));
    for (int p = 1; p < 4; ++p) {
        timed(c, st, "seg_counts", [&] {
            launch_seg_counts(prm.shape, ov_grid, st, d_keys, keysB, p, sp2, tinfo, counts, shifts + (p - 1), true);
        });
        timed(c, st, "seg_scan", [&] {
            hipLaunchKernelGGL(k_seg_scan, dim3(list_grid2), dim3(SEG_SCAN_GROUPS * SEG_SCAN_PAIRS), 0, st, pairs, sp2,
                               tile02, large2, counts, offsets, tile, INT64_MAX);
        });
        timed(c, st, "seg_scatter", [&] {
            launch_seg_scatter(prm.shape, prm.rank, ov_grid, st, d_keys, keysB, p, sp2, tinfo, offsets, nullptr,
                               nullptr, shifts + (p - 1), true);
        });
    }
}

// Radix top-k: the m largest keys of d_in[0, n) into d_out[0, m), in
// descending order (desc) or ascending; with d_rest, the other n - m keys go
// there in no particular order.  Scratch: [TopkState | cand 0 | cand 1].
// One-workgroup sort of the selection: counting ranks for small sets, an LSD
// radix sort in LDS for larger ones.
template <typename K, int FLOATK>
void launch_topk_sort(const K *src, uint32_t len, K *dst, uint32_t m, int mode, const TopkState *ts, uint32_t n,
                      int desc, hipStream_t st)
{
    if (len <= uint32_t(TOPK_RANK_SORT)) {
        uint32_t t = 64;
        while (t < len) t <<= 1;
        hipLaunchKernelGGL((k_topk_rank_sort<K, FLOATK>), dim3(1), dim3(t), 0, st, src, len, dst, m, mode, ts, n, desc);
    } else if (len <= uint32_t(256 * TOPK_SORT_ITEMS)) {
        hipLaunchKernelGGL((k_topk_radix_sort<K, FLOATK, 256>), dim3(1), dim3(256), 0, st, src, len, dst, m, mode, ts,
                           n, desc);
    } else {
        hipLaunchKernelGGL((k_topk_radix_sort<K, FLOATK, TOPK_SORT_BLOCK>), dim3(1), dim3(TOPK_SORT_BLOCK), 0, st, src,
                           len, dst, m, mode, ts, n, desc);
    }
}

template <typename K, int FLOATK>
void run_topk(hpcs_ctx *c, const K *d_in, int64_t n, int64_t m, K *d_out, K *d_rest, bool desc, hipStream_t st)
{
    using LV = TopkLevels<K>;
    if (n < 0 || m < 0 || m > n) fail(HPCS_E_INVALID_ARG, "k must lie in [0, n]");
    if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
    if (n > 0 && !d_in) fail(HPCS_E_INVALID_ARG, "keys is null");
    if (m > 0 && !d_out) fail(HPCS_E_INVALID_ARG, "output is null");
    if (reinterpret_cast<uintptr_t>(d_in) % sizeof(K) || reinterpret_cast<uintptr_t>(d_out) % sizeof(K) ||
        reinterpret_cast<uintptr_t>(d_rest) % sizeof(K))
        fail(HPCS_E_INVALID_ARG, "pointer is misaligned");
    if (m == 0) {
        if (d_rest && n > 0)
            hip_check(hipMemcpyAsync(d_rest, d_in, size_t(n) * sizeof(K), hipMemcpyDeviceToDevice, st), "copy rest");
        return;
    }
    const uint32_t un = uint32_t(n), um = uint32_t(m);
    const int idesc = desc ? 1 : 0;
    if (!d_rest && n <= TOPK_BLOCK_SORT) {
        // the whole input fits one workgroup: sort it there
        timed(c, st, "topk_sort", [&] { launch_topk_sort<K, FLOATK>(d_in, un, d_out, um, 0, nullptr, un, idesc, st); });
        return;
    }
    const size_t off_c0 = align_up(sizeof(TopkState));
    const size_t off_c1 = align_up(off_c0 + size_t(n) * sizeof(K));
    const size_t need = align_up(off_c1 + size_t(n) * sizeof(K));
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, need));
    TopkState *ts = reinterpret_cast<TopkState *>(base);
    K *cand[2] = {reinterpret_cast<K *>(base + off_c0), reinterpret_cast<K *>(base + off_c1)};
    const int vec = (reinterpret_cast<uintptr_t>(d_in) % 16 == 0) ? 1 : 0;
    constexpr int TILE = TopkTile<K>::KEYS;
    // one resident wave of workgroups (a second, partial round would leave
    // most CUs idle at the end of the grid-stride loop)
    static const int occ = [] { // resident compaction workgroups per CU, initialised once
        int o = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, k_topk_compact<K, FLOATK>, TOPK_BLOCK, 0) != hipSuccess ||
            o < 1)
            o = 1;
        return o;
    }();
    const int cgrid = grid_for(uint64_t(n), TILE, c->num_cus * occ);
    // large inputs without `rest`: sample, then keep the keys of the top bins
    // in one read (level 0 then descends over those alone)
    const int sampled = (!d_rest && un >= TOPK_FILTER_MIN && uint64_t(m) * 64u <= uint64_t(n)) ? 1 : 0;
    launch_fill(c, ts, sizeof(TopkState), st);
    if (sampled) {
        timed(c, st, "topk_sample", [&] {
            hipLaunchKernelGGL((k_topk_sample<K, FLOATK>), dim3(TOPK_SAMPLE_BLOCKS), dim3(TOPK_BLOCK), 0, st, d_in,
                               un, ts);
        });
        timed(c, st, "topk_filter", [&] {
            hipLaunchKernelGGL((k_topk_filter<K, FLOATK>), dim3(cgrid), dim3(TOPK_BLOCK), 0, st, d_in, un, vec, ts,
                               cand[0], um);
        });
    }
    timed(c, st, "topk_hist", [&] {
        hipLaunchKernelGGL((k_topk_hist<K, FLOATK>),
                           dim3(grid_for(uint64_t(n), TOPK_HIST_BLOCK * TopkTile<K>::ITEMS, c->num_cus * 2)),
                           dim3(TOPK_HIST_BLOCK), 0, st, d_in, un, vec, ts, sampled, um);
    });
    for (int l = 0; l < LV::LEVELS; ++l) {
        timed(c, st, "topk_compact", [&] {
            hipLaunchKernelGGL((k_topk_compact<K, FLOATK>), dim3(cgrid), dim3(TOPK_BLOCK), 0, st, d_in, un, vec,
                               cand[l & 1], cand[(l + 1) & 1], l, ts, d_out, d_rest, um, idesc, sampled);
        });
    }
    const uint64_t span = std::max<uint64_t>(uint64_t(m), uint64_t(n - m));
    if (m <= TOPK_BLOCK_SORT) {
        if (d_rest)
            timed(c, st, "topk_finish", [&] {
                hipLaunchKernelGGL((k_topk_finish<K, FLOATK>), dim3(grid_for(span, 256, c->num_cus * 4)), dim3(256),
                                   0, st, ts, static_cast<K *>(nullptr), d_rest, um, un, idesc);
            });
        timed(c, st, "topk_sort", [&] { launch_topk_sort<K, FLOATK>(d_out, um, d_out, um, 1, ts, un, idesc, st); });
        return;
    }
    timed(c, st, "topk_finish", [&] {
        hipLaunchKernelGGL((k_topk_finish<K, FLOATK>), dim3(grid_for(span, 256, c->num_cus * 4)), dim3(256), 0, st,
                           ts, d_out, d_rest, um, un, idesc);
    });
    run_sort<K, false, false>(c, d_out, nullptr, m, st);
    timed(c, st, "topk_values", [&] {
        hipLaunchKernelGGL((k_topk_values<K, FLOATK>), dim3(grid_for(uint64_t(m), 256, c->num_cus * 8)), dim3(256),
                           0, st, d_out, uint64_t(m), idesc);
    });
}

// SelectTopK(uint[], start, length, k) in place: a[k, start + length) gets the
// largest start + length - k keys of the range in ascending order, a[start, k)
// the others in no particular order.
template <typename K, int FLOATK>
void host_select_topk(hpcs_ctx *c, K *a, int64_t start, int64_t length, int64_t k)
{
    if (!a) fail(HPCS_E_INVALID_ARG, "arrayToBeSelected is null");
    if (start < 0 || length <= 0) fail(HPCS_E_INVALID_ARG, "start or length is invalid");
    if (k < start || k >= start + length) fail(HPCS_E_INVALID_ARG, "k must lie in [start, start + length)");
    if (length >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "length must be below 2^31");
    const int64_t m = start + length - k, r = length - m;
    const size_t off_out = align_up(size_t(length) * sizeof(K));
    const size_t off_rest = align_up(off_out + size_t(m) * sizeof(K));
    char *io = static_cast<char *>(ensure(c->io, c->io_cap, off_rest + size_t(r) * sizeof(K) + sizeof(K)));
    K *d_in = reinterpret_cast<K *>(io), *d_out = reinterpret_cast<K *>(io + off_out);
    K *d_rest = reinterpret_cast<K *>(io + off_rest);
    hipStream_t st = c->stream;
    hip_check(hipMemcpyAsync(d_in, a + start, size_t(length) * sizeof(K), hipMemcpyHostToDevice, st), "H2D");
    run_topk<K, FLOATK>(c, d_in, length, m, d_out, d_rest, false, st);
    if (r > 0) hip_check(hipMemcpyAsync(a + start, d_rest, size_t(r) * sizeof(K), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipMemcpyAsync(a + k, d_out, size_t(m) * sizeof(K), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// Host arrays: out[0, k) = the k largest keys of keys[0, n), descending.
template <typename K, int FLOATK>
void host_topk(hpcs_ctx *c, const K *keys, int64_t n, int64_t k, K *out)
{
    if (n < 0 || k < 0 || k > n) fail(HPCS_E_INVALID_ARG, "k must lie in [0, n]");
    if (n > 0 && !keys) fail(HPCS_E_INVALID_ARG, "keys is null");
    if (k > 0 && !out) fail(HPCS_E_INVALID_ARG, "output is null");
    if (k == 0) return;
    if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
    const size_t off_out = align_up(size_t(n) * sizeof(K));
    char *io = static_cast<char *>(ensure(c->io, c->io_cap, off_out + size_t(k) * sizeof(K)));
    K *d_in = reinterpret_cast<K *>(io), *d_out = reinterpret_cast<K *>(io + off_out);
    hipStream_t st = c->stream;
    hip_check(hipMemcpyAsync(d_in, keys, size_t(n) * sizeof(K), hipMemcpyHostToDevice, st), "H2D");
    run_topk<K, FLOATK>(c, d_in, n, k, d_out, nullptr, true, st);
    hip_check(hipMemcpyAsync(out, d_out, size_t(k) * sizeof(K), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// k_gather_records access width: 2 for 16-byte vectors (16-byte aligned
// records whose size is a multiple of 16), 1 for words, 0 for bytes.
int gather_mode(const void *recs, int32_t rec_size)
{
    const uintptr_t a = reinterpret_cast<uintptr_t>(recs);
    if ((rec_size & 15) == 0 && (a & 15) == 0) return 2;
    return ((rec_size & 3) == 0 && (a & 3) == 0) ? 1 : 0;
}

template <typename K>
void run_records(hpcs_ctx *c, void *d_recs, int64_t n, int32_t rec_size, int32_t key_offset, int32_t key_kind,
                 hipStream_t st)
{
    // The context's work buffer holds [keys n][idx n][copy of the records];
    // the sort's own scratch is the context's scratch.  The records are
    // gathered from the copy straight into place.
    const size_t off_idx = align_up(size_t(n) * sizeof(K));
    const size_t off_copy = align_up(off_idx + size_t(n) * sizeof(uint32_t));
    char *w = static_cast<char *>(ensure(c->work, c->work_cap, off_copy + size_t(n) * size_t(rec_size)));
    K *keys = reinterpret_cast<K *>(w);
    uint32_t *idx = reinterpret_cast<uint32_t *>(w + off_idx);
    unsigned char *copy = reinterpret_cast<unsigned char *>(w + off_copy);
    const int words = gather_mode(d_recs, rec_size);
    const int grid = grid_for(uint64_t(n), 256, c->num_cus * 8);
    // 16-byte vector records are copied by the extraction, which reads their
    // lines anyway; other records by a device copy
    if (words != 2)
        hip_check(hipMemcpyAsync(copy, d_recs, size_t(n) * size_t(rec_size), hipMemcpyDeviceToDevice, st),
                  "hipMemcpyAsync records");
    timed(c, st, "extract", [&] {
        hipLaunchKernelGGL((k_extract_keys<K>), dim3(grid), dim3(256), 0, st,
                           static_cast<const unsigned char *>(d_recs), uint64_t(n), rec_size, key_offset, key_kind,
                           keys, idx, words == 2 ? reinterpret_cast<u32x4 *>(copy) : static_cast<u32x4 *>(nullptr));
    });
    hip_check(hipGetLastError(), "k_extract_keys launch");
    run_sort<K, true, false>(c, keys, idx, n, st);
    timed(c, st, "gather", [&] {
        hipLaunchKernelGGL(k_gather_records,
                           dim3(grid_for(uint64_t(n) * uint64_t(rec_size) / 4 + 1, 256, c->num_cus * 8)), dim3(256),
                           0, st, copy, static_cast<unsigned char *>(d_recs), idx, uint64_t(n), rec_size, words);
    });
    hip_check(hipGetLastError(), "k_gather_records launch");
}

// Items sorted by a separate key array (the keys are sorted alongside):
// (key, index) pairs through the stable sort, then the items gathered by
// index from a copy (the context's work buffer) back into place.
template <typename K>
void run_keys_items(hpcs_ctx *c, K *d_keys, void *d_items, int64_t n, int32_t item_size, hipStream_t st)
{
    const size_t off_copy = align_up(size_t(n) * sizeof(uint32_t));
    char *w = static_cast<char *>(ensure(c->work, c->work_cap, off_copy + size_t(n) * size_t(item_size)));
    uint32_t *idx = reinterpret_cast<uint32_t *>(w);
    unsigned char *copy = reinterpret_cast<unsigned char *>(w + off_copy);
    hip_check(hipMemcpyAsync(copy, d_items, size_t(n) * size_t(item_size), hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync items");
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(uint64_t(n), 256, c->num_cus * 8)), dim3(256), 0, st, idx,
                       uint64_t(n));
    hip_check(hipGetLastError(), "k_iota_u32 launch");
    run_sort<K, true, KIND_UINT>(c, d_keys, idx, n, st);
    const int words = gather_mode(d_items, item_size);
    hipLaunchKernelGGL(k_gather_records,
                       dim3(grid_for(uint64_t(n) * uint64_t(item_size) / 4 + 1, 256, c->num_cus * 8)), dim3(256), 0,
                       st, copy, static_cast<unsigned char *>(d_items), idx, uint64_t(n), item_size, words);
    hip_check(hipGetLastError(), "k_gather_records launch");
}

void check_keys_items_args(int64_t n, int32_t key_bits, int32_t item_size)
{
    if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
    if (key_bits != 32 && key_bits != 64) fail(HPCS_E_INVALID_ARG, "key_bits must be 32 or 64");
    if (item_size <= 0) fail(HPCS_E_INVALID_ARG, "item_size must be positive");
    if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
}

void dev_keys_items(hpcs_ctx *c, void *d_keys, int32_t key_bits, void *d_items, int32_t item_size, int64_t n,
                    hipStream_t st)
{
    check_keys_items_args(n, key_bits, item_size);
    if (n <= 1) return;
    if (!d_keys || !d_items) fail(HPCS_E_INVALID_ARG, "keys or items is null");
    if (key_bits == 32) run_keys_items<uint32_t>(c, static_cast<uint32_t *>(d_keys), d_items, n, item_size, st);
    else run_keys_items<uint64_t>(c, static_cast<uint64_t *>(d_keys), d_items, n, item_size, st);
}

void check_records_args(int64_t n, int32_t rec_size, int32_t key_offset, int32_t key_bits, int32_t key_kind)
{
    if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
    if (key_bits != 32 && key_bits != 64) fail(HPCS_E_INVALID_ARG, "key_bits must be 32 or 64");
    if (key_kind != HPCS_KEY_UINT && key_kind != HPCS_KEY_FLOAT && key_kind != HPCS_KEY_INT)
        fail(HPCS_E_INVALID_ARG, "unknown key_kind");
    if (rec_size <= 0 || key_offset < 0 || key_offset + key_bits / 8 > rec_size)
        fail(HPCS_E_INVALID_ARG, "key does not fit inside the record");
    if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
}

void dev_records(hpcs_ctx *c, void *d_recs, int64_t n, int32_t rec_size, int32_t key_offset, int32_t key_bits,
                 int32_t key_kind, hipStream_t st)
{
    check_records_args(n, rec_size, key_offset, key_bits, key_kind);
    if (n <= 1) return;
    if (!d_recs) fail(HPCS_E_INVALID_ARG, "records is null");
    if (key_bits == 32)
        run_records<uint32_t>(c, d_recs, n, rec_size, key_offset, key_kind, st);
    else
        run_records<uint64_t>(c, d_recs, n, rec_size, key_offset, key_kind, st);
}

template <typename K, int FLOATK>
void dev_histogram(hpcs_ctx *c, const K *d_keys, int64_t n, uint32_t *d_counts, hipStream_t st)
{
    if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
    if (!d_counts) fail(HPCS_E_INVALID_ARG, "counts is null");
    if (n > 0 && !d_keys) fail(HPCS_E_INVALID_ARG, "keys is null");
    if (n >= (int64_t(1) << 32)) fail(HPCS_E_UNSUPPORTED, "histogram counts are 32-bit");
    launch_fill(c, d_counts, sizeof(K) * RADIX * sizeof(uint32_t), st);
    if (n > 0) launch_histogram<K, FLOATK>(c, d_keys, uint64_t(n), d_counts, st);
}

template <typename K>
int nbit_digits(int bits)
{
    if (bits < 1 || bits > 16) fail(HPCS_E_INVALID_ARG, "bits per digit must lie in [1, 16]");
    return (int(8 * sizeof(K)) + bits - 1) / bits;
}

template <typename K>
void dev_histogram_nbit(hpcs_ctx *c, const K *d_keys, int64_t n, int bits, uint32_t *d_counts, hipStream_t st)
{
    const int digits = nbit_digits<K>(bits);
    if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
    if (!d_counts) fail(HPCS_E_INVALID_ARG, "counts is null");
    if (n > 0 && !d_keys) fail(HPCS_E_INVALID_ARG, "keys is null");
    if (n >= (int64_t(1) << 32)) fail(HPCS_E_UNSUPPORTED, "histogram counts are 32-bit");
    const size_t table = size_t(digits) << bits;
    launch_fill(c, d_counts, table * sizeof(uint32_t), st);
    if (n == 0) return;
    const size_t lds = table * sizeof(uint32_t);
    const int in_lds = lds <= 64 * 1024 ? 1 : 0;
    timed(c, st, "histogram_nbit", [&] {
        hipLaunchKernelGGL((k_histogram_nbit<K>), dim3(grid_for(uint64_t(n), 256, c->num_cus * (in_lds ? 2 : 8))),
                           dim3(256), in_lds ? lds : 0, st, d_keys, uint64_t(n), bits, digits, in_lds, d_counts);
    });
}

template <typename K>
void host_histogram_nbit(hpcs_ctx *c, const K *a, int64_t l, int64_t r, int bits, int32_t *counts)
{
    const int digits = nbit_digits<K>(bits);
    if (!a) fail(HPCS_E_INVALID_ARG, "inArray is null");
    if (!counts) fail(HPCS_E_INVALID_ARG, "counts is null");
    if (l < 0) fail(HPCS_E_INVALID_ARG, "l must be non-negative");
    const int64_t n = r >= l ? r - l + 1 : 0;
    if (n >= (int64_t(1) << 31)) fail(HPCS_E_UNSUPPORTED, "int32 counts limit the range to 2^31 - 1 keys");
    const size_t table = size_t(digits) << bits;
    const size_t hoff = align_up(size_t(n) * sizeof(K));
    char *io = static_cast<char *>(ensure(c->io, c->io_cap, hoff + table * sizeof(uint32_t)));
    K *d_keys = reinterpret_cast<K *>(io);
    uint32_t *d_counts = reinterpret_cast<uint32_t *>(io + hoff);
    hipStream_t st = c->stream;
    if (n > 0) hip_check(hipMemcpyAsync(d_keys, a + l, size_t(n) * sizeof(K), hipMemcpyHostToDevice, st), "H2D");
    dev_histogram_nbit<K>(c, d_keys, n, bits, d_counts, st);
    hip_check(hipMemcpyAsync(counts, d_counts, table * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

// Device-resident calls run on the caller's stream; NULL is HIP's null stream.
hipStream_t pick_stream(hpcs_ctx *, void *stream) { return static_cast<hipStream_t>(stream); }

// ---- graph cache ----
constexpr size_t GRAPH_CACHE_MAX = 64;

// Destroy every cached graph; the device is drained first, since a graph
// launched earlier may still run.
void graph_cache_clear(hpcs_ctx *c)
{
    bool any = false;
    for (auto &e : c->graph_cache) any = any || e.exec != nullptr;
    if (any) hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    for (auto &e : c->graph_cache)
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
    c->graph_cache.clear();
}

// Run the launch sequence `f` on stream st, through the graph cache when it is
// on (and the context is not profiling, st is not the null stream and is not
// being captured by the caller).  The first call with a key runs `f`
// directly, which sizes the context's buffers; the next call with the same
// key, while those buffers are unchanged, captures `f` into a graph (so no
// allocation happens under capture) and launches it; later calls launch the
// graph.  Every input of the launch sequence is in the key or in the
// context's options, and changing an option clears the cache.
template <typename F>
void run_cached(hpcs_ctx *c, const hpcs_ctx::GraphEntry &key, hipStream_t st, F &&f)
{
    if (!c->graphs || c->profile || st == nullptr) return f();
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(st, &cs), "hipStreamIsCapturing");
    if (cs != hipStreamCaptureStatusNone) return f();
    auto &cache = c->graph_cache;
    size_t i = 0;
    while (i < cache.size() && !(cache[i].entry == key.entry && cache[i].p0 == key.p0 && cache[i].p1 == key.p1 &&
                                 cache[i].n == key.n && cache[i].aux == key.aux && cache[i].st == key.st))
        ++i;
    const bool same_buffers = i < cache.size() && cache[i].scratch == c->scratch && cache[i].io == c->io &&
                              cache[i].scratch_cap == c->scratch_cap && cache[i].io_cap == c->io_cap;
    if (i == cache.size() || !same_buffers) {
        // first sighting (or the buffers moved): run directly, then remember
        // the key with the buffers it used
        f();
        if (i < cache.size() && cache[i].exec) {
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            (void)hipGraphExecDestroy(cache[i].exec);
            cache[i].exec = nullptr;
        }
        if (i == cache.size()) {
            if (cache.size() >= GRAPH_CACHE_MAX) {
                // evict the least recently used entry, preferring one with no graph
                size_t v = 0;
                for (size_t j = 1; j < cache.size(); ++j) {
                    const bool jg = cache[j].exec != nullptr, vg = cache[v].exec != nullptr;
                    if ((!jg && vg) || (jg == vg && cache[j].last_use < cache[v].last_use)) v = j;
                }
                if (cache[v].exec) {
                    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
                    (void)hipGraphExecDestroy(cache[v].exec);
                }
                cache.erase(cache.begin() + std::ptrdiff_t(v));
            }
            cache.push_back(key);
            i = cache.size() - 1;
            cache[i].exec = nullptr;
        }
        cache[i].scratch = c->scratch;
        cache[i].io = c->io;
        cache[i].scratch_cap = c->scratch_cap;
        cache[i].io_cap = c->io_cap;
        cache[i].last_use = ++c->graph_clock;
        return;
    }
    auto &e = cache[i];
    e.last_use = ++c->graph_clock;
    if (!e.exec) {
        hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        hipGraph_t g = nullptr;
        try {
            f();
        } catch (...) {
            if (hipStreamEndCapture(st, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(st, &g), "hipStreamEndCapture");
        hipGraphExec_t ex = nullptr;
        const hipError_t err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        hip_check(err, "hipGraphInstantiate");
        e.exec = ex;
        ++c->graph_captures;
    }
    hip_check(hipGraphLaunch(e.exec, st), "hipGraphLaunch");
    ++c->graph_replays;
}

// A device-resident sort through the graph cache.
template <typename K, bool HAS_VALS, int KIND>
void dev_sort(hpcs_ctx *c, K *d_keys, uint32_t *d_vals, int64_t n, hipStream_t st)
{
    if (n <= 1) return run_sort<K, HAS_VALS, KIND>(c, d_keys, d_vals, n, st);
    const hpcs_ctx::GraphEntry key{int(sizeof(K)) * 16 + (HAS_VALS ? 8 : 0) + KIND, d_keys, d_vals, n, 0, st};
    run_cached(c, key, st, [&] { run_sort_limited<K, HAS_VALS, KIND>(c, d_keys, d_vals, n, st); });
}

// Byte histogram of the key inside each record (the key extracted into the
// context's scratch, then the plain histogram).
void dev_histogram_records(hpcs_ctx *c, const void *d_recs, int64_t n, int32_t rec_size, int32_t key_offset,
                           int32_t key_bits, uint32_t *d_counts, hipStream_t st)
{
    check_records_args(n, rec_size, key_offset, key_bits, HPCS_KEY_UINT);
    if (!d_counts) fail(HPCS_E_INVALID_ARG, "counts is null");
    if (n > 0 && !d_recs) fail(HPCS_E_INVALID_ARG, "records is null");
    const size_t ks = size_t(key_bits / 8);
    char *base = static_cast<char *>(ensure(c->scratch, c->scratch_cap, align_up(size_t(std::max<int64_t>(n, 1)) * ks)));
    const int grid = grid_for(uint64_t(std::max<int64_t>(n, 1)), 256, c->num_cus * 8);
    auto recs = static_cast<const unsigned char *>(d_recs);
    if (key_bits == 32) {
        if (n > 0)
            timed(c, st, "extract", [&] {
                hipLaunchKernelGGL((k_extract_keys<uint32_t>), dim3(grid), dim3(256), 0, st, recs, uint64_t(n), rec_size,
                                   key_offset, int(HPCS_KEY_UINT), reinterpret_cast<uint32_t *>(base),
                                   static_cast<uint32_t *>(nullptr), static_cast<u32x4 *>(nullptr));
            });
        dev_histogram<uint32_t, KIND_UINT>(c, reinterpret_cast<const uint32_t *>(base), n, d_counts, st);
    } else {
        if (n > 0)
            timed(c, st, "extract", [&] {
                hipLaunchKernelGGL((k_extract_keys<uint64_t>), dim3(grid), dim3(256), 0, st, recs, uint64_t(n), rec_size,
                                   key_offset, int(HPCS_KEY_UINT), reinterpret_cast<uint64_t *>(base),
                                   static_cast<uint32_t *>(nullptr), static_cast<u32x4 *>(nullptr));
            });
        dev_histogram<uint64_t, KIND_UINT>(c, reinterpret_cast<const uint64_t *>(base), n, d_counts, st);
    }
}

// Host-array call: copy in, sort on the device, copy out, wait.
template <typename K, bool HAS_VALS, int FLOATK>
void host_sort(hpcs_ctx *c, K *keys, uint32_t *vals, int64_t n)
{
    if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
    if (n > 0 && !keys) fail(HPCS_E_INVALID_ARG, "keys is null");
    if (HAS_VALS && n > 0 && !vals) fail(HPCS_E_INVALID_ARG, "values is null");
    if (n <= 1) return;
    if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
    const size_t kbytes = size_t(n) * sizeof(K);
    const size_t voff = align_up(kbytes);
    const size_t need = voff + (HAS_VALS ? size_t(n) * sizeof(uint32_t) : 0);
    char *io = static_cast<char *>(ensure(c->io, c->io_cap, need));
    K *d_keys = reinterpret_cast<K *>(io);
    uint32_t *d_vals = HAS_VALS ? reinterpret_cast<uint32_t *>(io + voff) : nullptr;
    hipStream_t st = c->stream;
    hip_check(hipMemcpyAsync(d_keys, keys, kbytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D");
    if (HAS_VALS) hip_check(hipMemcpyAsync(d_vals, vals, size_t(n) * 4, hipMemcpyHostToDevice, st), "hipMemcpyAsync H2D");
    run_sort_limited<K, HAS_VALS, FLOATK>(c, d_keys, d_vals, n, st);
    hip_check(hipMemcpyAsync(keys, d_keys, kbytes, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
    if (HAS_VALS) hip_check(hipMemcpyAsync(vals, d_vals, size_t(n) * 4, hipMemcpyDeviceToHost, st), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
}

template <typename K, int KIND = KIND_UINT>
void host_histogram(hpcs_ctx *c, const K *a, int64_t l, int64_t r, int32_t *counts)
{
    if (!a) fail(HPCS_E_INVALID_ARG, "inArray is null");
    if (!counts) fail(HPCS_E_INVALID_ARG, "counts is null");
    if (l < 0) fail(HPCS_E_INVALID_ARG, "l must be non-negative");
    const int64_t n = r >= l ? r - l + 1 : 0;
    if (n >= (int64_t(1) << 31)) fail(HPCS_E_UNSUPPORTED, "int32 counts limit the range to 2^31 - 1 keys");
    constexpr int P = int(sizeof(K));
    const size_t hoff = align_up(size_t(n) * sizeof(K));
    char *io = static_cast<char *>(ensure(c->io, c->io_cap, hoff + P * RADIX * sizeof(uint32_t)));
    K *d_keys = reinterpret_cast<K *>(io);
    uint32_t *d_counts = reinterpret_cast<uint32_t *>(io + hoff);
    hipStream_t st = c->stream;
    if (n > 0) hip_check(hipMemcpyAsync(d_keys, a + l, size_t(n) * sizeof(K), hipMemcpyHostToDevice, st), "H2D");
    dev_histogram<K, KIND>(c, d_keys, n, d_counts, st);
    uint32_t tmp[8 * RADIX];
    hip_check(hipMemcpyAsync(tmp, d_counts, P * RADIX * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "D2H");
    hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (int i = 0; i < P * RADIX; ++i) counts[i] = int32_t(tmp[i]);
}

#if HPCS_IN_PART(0)
// The device address of `scratch` on `device` (current), resolved once per
// device and kept; null (with the reason in `why`) if the runtime gives none,
// which fails only the no-allocation entry point, not the context.
void *noalloc_scratch_address(int device, std::string &why)
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it != addr.end()) return it->second;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError(); // leave no sticky error for later launches
        why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return nullptr;
    }
    addr[device] = p;
    return p;
}
#endif

template <typename F>
int32_t guarded_host(F &&f)
{
    try {
        f();
        return HPCS_OK;
    } catch (const Err &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host out of memory";
        return HPCS_E_OOM;
    } catch (...) {
        g_last_error = "unexpected C++ exception";
        return HPCS_E_HIP;
    }
}

// The same for calls on a context: the context's lock and device are held.
template <typename F>
int32_t guarded(hpcs_ctx *c, F &&f)
{
    if (!c) {
        g_last_error = "context is null";
        return HPCS_E_INVALID_ARG;
    }
    try {
        std::lock_guard<std::mutex> lock(c->mu);
        DeviceGuard dg(c->device);
        f();
        return HPCS_OK;
    } catch (const Err &e) {
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host out of memory";
        return HPCS_E_OOM;
    } catch (...) {
        g_last_error = "unexpected C++ exception";
        return HPCS_E_HIP;
    }
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_version(void) { return 100; }

HPCS_API const char *hpcs_last_error(void) { return g_last_error.c_str(); }

HPCS_API int32_t hpcs_ctx_create(int32_t device, hpcs_ctx **out)
{
    if (!out) {
        g_last_error = "out is null";
        return HPCS_E_INVALID_ARG;
    }
    *out = nullptr;
    hpcs_ctx *c = nullptr;
    try {
        int count = 0;
        hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
        if (device < 0 || device >= count) fail(HPCS_E_INVALID_ARG, "no such HIP device");
        c = new hpcs_ctx();
        c->device = device;
        DeviceGuard dg(device);
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        if (const char *r = std::getenv("HPCS_RANK")) {
            c->rank_mode = std::strcmp(r, "ballot") == 0 ? RANK_BALLOT
                           : std::strcmp(r, "mix2") == 0  ? RANK_MIX2
                           : std::strcmp(r, "mix4") == 0  ? RANK_MIX4
                                                          : RANK_LDSOR;
            c->set_rank = true;
        }
        if (const char *t = std::getenv("HPCS_TILE_SHAPE")) {
            const int v = std::atoi(t);
            c->tile_shape = (v >= 0 && v < NUM_TILE_SHAPES) ? v : 0;
            c->set_shape = true;
        }
        if (const char *tf = std::getenv("HPCS_TUNE_FILE")) {
            std::string text, err;
            if (!read_file(tf, text)) fail(HPCS_E_INVALID_ARG, std::string("cannot read HPCS_TUNE_FILE ") + tf);
            if (!parse_tuning(text, c->tune, err)) fail(HPCS_E_INVALID_ARG, "HPCS_TUNE_FILE: " + err);
        }
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        c->noalloc_scratch = noalloc_scratch_address(device, c->noalloc_error);
        *out = c;
        return HPCS_OK;
    } catch (const Err &e) {
        delete c;
        g_last_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        delete c;
        g_last_error = "host out of memory";
        return HPCS_E_OOM;
    }
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_ctx_trim(hpcs_ctx *ctx)
{
    return guarded(ctx, [&] {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        graph_cache_clear(ctx);
        if (ctx->scratch) hip_check(hipFree(ctx->scratch), "hipFree");
        if (ctx->io) hip_check(hipFree(ctx->io), "hipFree");
        if (ctx->work) hip_check(hipFree(ctx->work), "hipFree");
        ctx->scratch = ctx->io = ctx->work = nullptr;
        ctx->scratch_cap = ctx->io_cap = ctx->work_cap = 0;
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_ctx_destroy(hpcs_ctx *ctx)
{
    if (!ctx) return HPCS_OK;
    int32_t rc = hpcs_ctx_trim(ctx);
    try {
        DeviceGuard dg(ctx->device);
        for (auto &p : ctx->pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto e : ctx->ev_pool) (void)hipEventDestroy(e);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    } catch (const Err &) {
    }
    delete ctx;
    return rc;
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_ctx_set_option(hpcs_ctx *ctx, int32_t option, int64_t value)
{
    return guarded(ctx, [&] {
        graph_cache_clear(ctx); // cached launch sequences follow the options
        switch (option) {
        case HPCS_OPT_GRAPHS:
            if (value != 0 && value != 1) fail(HPCS_E_INVALID_ARG, "graphs must be 0 or 1");
            ctx->graphs = int(value);
            break;
        case HPCS_OPT_RANK:
            if (value < RANK_BALLOT || value > RANK_MIX4) fail(HPCS_E_INVALID_ARG, "rank mode must be 0, 1, 2 or 3");
            ctx->rank_mode = int(value);
            ctx->set_rank = true;
            break;
        case HPCS_OPT_TILE_SHAPE:
            if (value < 0 || value >= NUM_TILE_SHAPES) fail(HPCS_E_INVALID_ARG, "unknown tile shape");
            ctx->tile_shape = int(value);
            ctx->set_shape = true;
            break;
        case HPCS_OPT_LOOKBACK_WINDOW:
            if (value != 16 && value != 32 && value != 64) fail(HPCS_E_INVALID_ARG, "window must be 16, 32 or 64");
            ctx->lookback_window = int(value);
            break;
        case HPCS_OPT_PASS_ALGO:
            if (value != PASS_ONESWEEP && value != PASS_RTS && value != PASS_AUTO && value != PASS_BLOCK &&
                value != PASS_MSD && value != PASS_MERGE && value != PASS_HYBRID)
                fail(HPCS_E_INVALID_ARG, "pass algorithm must be 0 to 6");
            ctx->pass_algo = int(value);
            ctx->set_algo = value != PASS_AUTO;
            break;
        case HPCS_OPT_DIGIT_BITS:
            if (value != 8 && value != 11) fail(HPCS_E_INVALID_ARG, "digit bits must be 8 or 11");
            ctx->digit_bits = int(value);
            ctx->set_bits = true;
            break;
        case HPCS_OPT_DIGIT_BYTES:
            if (value < 0 || value > 2) fail(HPCS_E_INVALID_ARG, "digit bytes must be 0, 1 or 2");
            ctx->digit_bytes = int(value);
            break;
        case HPCS_OPT_SCRATCH_LIMIT:
            if (value < 0) fail(HPCS_E_INVALID_ARG, "scratch limit must be non-negative");
            ctx->scratch_limit = size_t(value);
            break;
        case HPCS_OPT_SEG_RANK:
            if (value < RANK_BALLOT || value > RANK_MIX4) fail(HPCS_E_INVALID_ARG, "rank mode must be 0, 1, 2 or 3");
            ctx->seg_rank = int(value);
            ctx->set_seg_rank = true;
            break;
        case HPCS_OPT_SEG_SHAPE:
            if (value < 0 || value >= NUM_TILE_SHAPES || !seg_shape_ok(int(value)))
                fail(HPCS_E_INVALID_ARG, "the tile shape does not serve the batched segments");
            ctx->seg_shape = int(value);
            ctx->set_seg_shape = true;
            break;
        case HPCS_OPT_SEG_SMALL:
            if (!seg_small_valid(value)) fail(HPCS_E_INVALID_ARG, "the LDS segment limit must be 4096 or 16384");
            ctx->seg_small = int(value);
            ctx->set_seg_small = true;
            break;
        case HPCS_OPT_SEG_MSD:
            if (value != 0 && (value <= SEG_MEDIUM || value >= (int64_t(1) << 32)))
                fail(HPCS_E_INVALID_ARG, "the MSD-first threshold must be 0 or in [16384, 2^32)");
            ctx->seg_msd = value;
            break;
        case HPCS_OPT_FALLBACK_ROUNDS:
            if (value < 0 || value > (1 << 20)) fail(HPCS_E_INVALID_ARG, "fallback rounds out of range");
            ctx->fallback_rounds = int(value);
            break;
        default: fail(HPCS_E_INVALID_ARG, "unknown option");
        }
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_profile_enable(hpcs_ctx *ctx, int32_t on)
{
    return guarded(ctx, [&] {
        harvest(ctx);
        ctx->stats.clear();
        ctx->profile = on != 0;
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_profile_read(hpcs_ctx *ctx, hpcs_kernel_stat *out, int32_t max, int32_t *count)
{
    return guarded(ctx, [&] {
        if (!count || (max > 0 && !out)) fail(HPCS_E_INVALID_ARG, "out/count is null");
        harvest(ctx);
        int32_t i = 0;
        for (const auto &kv : ctx->stats) {
            if (i >= max) break;
            std::memset(out[i].name, 0, sizeof(out[i].name));
            std::strncpy(out[i].name, kv.first.c_str(), sizeof(out[i].name) - 1);
            out[i].launches = kv.second.first;
            out[i].total_ms = kv.second.second;
            ++i;
        }
        *count = i;
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_graph_cache_stats(hpcs_ctx *ctx, int64_t *entries, int64_t *captures, int64_t *replays)
{
    return guarded(ctx, [&] {
        if (!entries || !captures || !replays) fail(HPCS_E_INVALID_ARG, "null output");
        int64_t g = 0;
        for (auto &e : ctx->graph_cache) g += e.exec != nullptr;
        *entries = g;
        *captures = ctx->graph_captures;
        *replays = ctx->graph_replays;
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_tile_shape_query(int32_t shape, int32_t key_bytes, int32_t with_values, int32_t *block,
                                       int32_t *items)
{
    return guarded_host([&] {
        if (!block || !items) fail(HPCS_E_INVALID_ARG, "block/items is null");
        if (key_bytes != 4 && key_bytes != 8) fail(HPCS_E_INVALID_ARG, "key_bytes must be 4 or 8");
        // the dispatch every pass launcher goes through
        auto report = [&](auto sc) {
            constexpr TileShape T = TILE_SHAPES[decltype(sc)::value];
            *block = T.block;
            *items = T.items;
        };
        if (key_bytes == 4 && with_values) with_shape<4, true>(shape, report);
        else if (key_bytes == 4) with_shape<4, false>(shape, report);
        else if (with_values) with_shape<8, true>(shape, report);
        else with_shape<8, false>(shape, report);
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_plan_query(hpcs_ctx *ctx, int32_t dtype, int64_t n, int32_t with_values, hpcs_plan_info *info)
{
    if (!info) {
        g_last_error = "info is null";
        return HPCS_E_INVALID_ARG;
    }
    (void)ctx;
    size_t ks;
    switch (dtype) {
    case HPCS_DTYPE_U32:
    case HPCS_DTYPE_F32: ks = 4; break;
    case HPCS_DTYPE_U64:
    case HPCS_DTYPE_F64: ks = 8; break;
    default: g_last_error = "unknown dtype"; return HPCS_E_INVALID_ARG;
    }
    if (n < 0) {
        g_last_error = "n must be non-negative";
        return HPCS_E_INVALID_ARG;
    }
    const SortParams prm = ctx ? resolve_params(ctx, int(ks), with_values != 0, n) : SortParams{PASS_AUTO, 0, 0, 8};
    if (prm.bits == 11) {
        const int passes = ks == 4 ? d11_passes<uint32_t>() : d11_passes<uint64_t>();
        const Layout11 L = make_layout11(n, ks);
        info->digit_bits = D11_BITS;
        info->passes = passes;
        info->tile_keys = D11_TILE;
        info->block_threads = D11_BLOCK;
        info->tiles = L.tiles;
        info->ideal_bytes = n * (int64_t(ks) + passes * 2 * int64_t(ks));
        info->scratch_bytes = int64_t(L.total);
        return HPCS_OK;
    }
    const int passes = int(ks * 8 / RADIX_BITS);
    if (prm.algo == PASS_MERGE) {
        // one read and one write per key for the runs, then per merge pass
        const int64_t run = ks == 4 ? 4096 : 2048;
        int merges = 0;
        while ((run << merges) < n) ++merges;
        info->digit_bits = RADIX_BITS;
        info->passes = passes;
        info->block_threads = 256;
        info->tile_keys = run;
        info->tiles = (n + run - 1) / run;
        info->ideal_bytes = n * int64_t(1 + merges) * 2 * (int64_t(ks) + (with_values ? 4 : 0));
        info->scratch_bytes = int64_t(merge_sort_scratch_bytes(n, ks, with_values != 0));
        return HPCS_OK;
    }
    if (prm.algo == PASS_BLOCK) {
        const int items = ks == 4 ? 16 : 8;
        info->digit_bits = RADIX_BITS;
        info->passes = passes;
        info->block_threads = n <= 256 * items ? 256 : 1024;
        info->tile_keys = info->block_threads * items;
        info->tiles = 1;
        info->ideal_bytes = n * (int64_t(ks) + passes * 2 * (int64_t(ks) + (with_values ? 4 : 0)));
        info->scratch_bytes = 0;
        return HPCS_OK;
    }
    const int shape = prm.shape;
    const int tile_keys = TILE_SHAPES[shape].block * TILE_SHAPES[shape].items;
    const Layout L =
        make_layout(n, ks, with_values != 0, tile_keys,
                    prm.algo != PASS_MSD && (ctx ? digit_bytes_on(ctx, ks, with_values != 0) : ks == 8 && with_values));
    if (prm.algo == PASS_MSD) {
        // two reads (OR/AND, split histogram), the split (read + write), the
        // bucket sorts (read + write; buckets within a tile)
        info->digit_bits = RADIX_BITS;
        info->passes = 2;
        info->tile_keys = tile_keys;
        info->block_threads = TILE_SHAPES[shape].block;
        info->tiles = L.tiles;
        info->ideal_bytes = n * int64_t(ks) * 6;
        info->scratch_bytes = int64_t(L.total);
        return HPCS_OK;
    }
    const int64_t per_key = int64_t(ks) + passes * 2 * (int64_t(ks) + (with_values ? 4 : 0));
    info->digit_bits = RADIX_BITS;
    info->passes = passes;
    info->tile_keys = tile_keys;
    info->block_threads = TILE_SHAPES[shape].block;
    info->tiles = L.tiles;
    // the LSD form's ideal (one counting read, a read and a write per pass),
    // also for the MSD-first form, which moves fewer bytes
    info->ideal_bytes = n * per_key;
    info->scratch_bytes = int64_t(L.total);
    if (prm.algo == PASS_HYBRID && dtype == HPCS_DTYPE_U32) info->scratch_bytes = int64_t(make_hyb_layout(n, shape).total);
    return HPCS_OK;
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_ctx_reserve(hpcs_ctx *ctx, int32_t dtype, int64_t n)
{
    return guarded(ctx, [&] {
        size_t ks = 0;
        switch (dtype) {
        case HPCS_DTYPE_U32:
        case HPCS_DTYPE_F32: ks = 4; break;
        case HPCS_DTYPE_U64:
        case HPCS_DTYPE_F64: ks = 8; break;
        default: fail(HPCS_E_INVALID_ARG, "unknown dtype");
        }
        if (n < 0) fail(HPCS_E_INVALID_ARG, "n must be non-negative");
        if (n >= MAX_KEYS) fail(HPCS_E_UNSUPPORTED, "n must be below 2^31");
        size_t need = 0;
        for (int shape = 0; shape < NUM_TILE_SHAPES; ++shape) {
            if (!shape_serves(shape, int(ks), false)) continue;
            const int tk = shape_tile_keys(shape);
            need = std::max(need, make_layout(n, ks, false, tk).total);
            if (ks == 4 && shape_serves(shape, 4, false) && n > BLOCK_SORT_MAX4 && n <= HYB_MAX_KEYS)
                need = std::max(need, make_hyb_layout(n, shape).total); // the MSD-first form
        }
        if (n < LOOKBACK_MAX_KEYS) need = std::max(need, make_layout11(n, ks).total);
        const size_t topk = align_up(align_up(align_up(sizeof(TopkState)) + size_t(n) * ks) + size_t(n) * ks);
        need = std::max(need, topk);
        ensure(ctx->scratch, ctx->scratch_cap, need);
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_ctx_reserve_segments(hpcs_ctx *ctx, int64_t n, int32_t nseg)
{
    return guarded(ctx, [&] {
        if (n < 0 || nseg < 0) fail(HPCS_E_INVALID_ARG, "n and nseg must be non-negative");
        if (n >= (int64_t(1) << 32)) fail(HPCS_E_UNSUPPORTED, "segments must lie below 2^32 keys");
        // the smallest tile bounds the scratch of every call of up to n keys,
        // whatever tile shape its size class takes (with the MSD-first
        // form's overflow plan when the context has a threshold)
        if (nseg > 0 && n > 1)
            ensure(ctx->scratch, ctx->scratch_cap, seg_layout(n, nseg, SEG_MIN_TILE, ctx->seg_msd).total);
    });
}
#endif

// host arrays
#if HPCS_IN_PART(1)
HPCS_API int32_t hpcs_sort_radix_u32(hpcs_ctx *ctx, uint32_t *keys, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint32_t, false, false>(ctx, keys, nullptr, n); });
}
#endif
#if HPCS_IN_PART(1)
HPCS_API int32_t hpcs_sort_radix_range_u32(hpcs_ctx *ctx, uint32_t *keys, int64_t start, int64_t length)
{
    return guarded(ctx, [&] {
        if (start < 0 || length < 0) fail(HPCS_E_INVALID_ARG, "start and length must be non-negative");
        if (length > 0 && !keys) fail(HPCS_E_INVALID_ARG, "keys is null");
        host_sort<uint32_t, false, false>(ctx, keys ? keys + start : keys, nullptr, length);
    });
}
#endif
#if HPCS_IN_PART(2)
HPCS_API int32_t hpcs_sort_radix_u64(hpcs_ctx *ctx, uint64_t *keys, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint64_t, false, false>(ctx, keys, nullptr, n); });
}
#endif
#if HPCS_IN_PART(5)
HPCS_API int32_t hpcs_sort_radix_f32(hpcs_ctx *ctx, float *keys, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint32_t, false, true>(ctx, reinterpret_cast<uint32_t *>(keys), nullptr, n); });
}
#endif
#if HPCS_IN_PART(2)
HPCS_API int32_t hpcs_sort_radix_f64(hpcs_ctx *ctx, double *keys, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint64_t, false, true>(ctx, reinterpret_cast<uint64_t *>(keys), nullptr, n); });
}
#endif
#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_sort_radix_pairs_u64_u32(hpcs_ctx *ctx, uint64_t *keys, uint32_t *vals, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint64_t, true, false>(ctx, keys, vals, n); });
}
#endif
#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_sort_radix_pairs_u32_u32(hpcs_ctx *ctx, uint32_t *keys, uint32_t *vals, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint32_t, true, false>(ctx, keys, vals, n); });
}
#endif
#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_sort_radix_records(hpcs_ctx *ctx, void *recs, int64_t n, int32_t rec_size, int32_t key_offset,
                                         int32_t key_bits, int32_t key_kind)
{
    return guarded(ctx, [&] {
        check_records_args(n, rec_size, key_offset, key_bits, key_kind);
        if (n > 0 && !recs) fail(HPCS_E_INVALID_ARG, "records is null");
        if (n <= 1) return;
        const size_t bytes = size_t(n) * size_t(rec_size);
        void *d = ensure(ctx->io, ctx->io_cap, bytes);
        hip_check(hipMemcpyAsync(d, recs, bytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
        dev_records(ctx, d, n, rec_size, key_offset, key_bits, key_kind, ctx->stream);
        hip_check(hipMemcpyAsync(recs, d, bytes, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}
#endif
#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_byte_components_u32(hpcs_ctx *ctx, const uint32_t *a, int64_t l, int64_t r,
                                                    int32_t *counts)
{
    return guarded(ctx, [&] { host_histogram<uint32_t>(ctx, a, l, r, counts); });
}
#endif
#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_byte_components_u64(hpcs_ctx *ctx, const uint64_t *a, int64_t l, int64_t r,
                                                    int32_t *counts)
{
    return guarded(ctx, [&] { host_histogram<uint64_t>(ctx, a, l, r, counts); });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_sort_segments_u32(hpcs_ctx *ctx, uint32_t *keys, const int64_t *seg_begin, int32_t nseg)
{
    return guarded(ctx, [&] {
        if (nseg < 0) fail(HPCS_E_INVALID_ARG, "nseg must be non-negative");
        if (nseg == 0) return;
        if (!seg_begin) fail(HPCS_E_INVALID_ARG, "seg_begin is null");
        if (seg_begin[0] < 0) fail(HPCS_E_INVALID_ARG, "segments must start at or after index 0");
        for (int32_t s = 0; s < nseg; ++s)
            if (seg_begin[s + 1] < seg_begin[s]) fail(HPCS_E_INVALID_ARG, "segment offsets must be non-decreasing");
        const int64_t n = seg_begin[nseg];
        if (n > 0 && !keys) fail(HPCS_E_INVALID_ARG, "keys is null");
        if (n <= 1) return;
        const size_t kbytes = size_t(n) * 4, boff = align_up(kbytes);
        char *io = static_cast<char *>(ensure(ctx->io, ctx->io_cap, boff + size_t(nseg + 1) * 8));
        uint32_t *d_keys = reinterpret_cast<uint32_t *>(io);
        int64_t *d_begin = reinterpret_cast<int64_t *>(io + boff);
        hip_check(hipMemcpyAsync(d_keys, keys, kbytes, hipMemcpyHostToDevice, ctx->stream), "H2D");
        hip_check(hipMemcpyAsync(d_begin, seg_begin, size_t(nseg + 1) * 8, hipMemcpyHostToDevice, ctx->stream), "H2D");
        run_segments(ctx, d_keys, n, d_begin, nseg, ctx->stream);
        hip_check(hipMemcpyAsync(keys, d_keys, kbytes, hipMemcpyDeviceToHost, ctx->stream), "D2H");
        hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    });
}
#endif

// device-resident
#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_sort_segments_u32(hpcs_ctx *ctx, uint32_t *d_keys, int64_t n, const int64_t *d_seg_begin,
                                            int32_t nseg, void *stream)
{
    return guarded(ctx, [&] {
        hipStream_t st = pick_stream(ctx, stream);
        if (n <= 1 || nseg <= 0) return run_segments(ctx, d_keys, n, d_seg_begin, nseg, st);
        const hpcs_ctx::GraphEntry key{1, d_keys, d_seg_begin, n, nseg, st};
        run_cached(ctx, key, st, [&] { run_segments(ctx, d_keys, n, d_seg_begin, nseg, st); });
    });
}
#endif

#if HPCS_IN_PART(1)
HPCS_API int32_t hpcs_dev_sort_radix_u32(hpcs_ctx *ctx, uint32_t *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] { dev_sort<uint32_t, false, false>(ctx, d_keys, nullptr, n, pick_stream(ctx, stream)); });
}
#endif
#if HPCS_IN_PART(1)
HPCS_API int32_t hpcs_dev_sort_radix_u32_noalloc(hpcs_ctx *ctx, uint32_t *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] {
        if (n > NOALLOC_MAX_KEYS)
            fail(HPCS_E_UNSUPPORTED, "hpcs_dev_sort_radix_u32_noalloc sorts at most 2^22 keys (n = " +
                                         std::to_string(n) + ")");
        if (!ctx->noalloc_scratch) fail(HPCS_E_HIP, "the static scratch has no device address: " + ctx->noalloc_error);
        // no event pool, graph cache or context scratch: nothing allocates
        const bool profile = ctx->profile;
        ctx->profile = false;
        const ExtScratch ext{ctx->noalloc_scratch, NOALLOC_SCRATCH_BYTES};
        try {
            run_sort<uint32_t, false, KIND_UINT>(ctx, d_keys, nullptr, n, pick_stream(ctx, stream), &ext);
        } catch (...) {
            ctx->profile = profile;
            throw;
        }
        ctx->profile = profile;
    });
}
#endif
#if HPCS_IN_PART(2)
HPCS_API int32_t hpcs_dev_sort_radix_u64(hpcs_ctx *ctx, uint64_t *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] { dev_sort<uint64_t, false, false>(ctx, d_keys, nullptr, n, pick_stream(ctx, stream)); });
}
#endif
#if HPCS_IN_PART(5)
HPCS_API int32_t hpcs_dev_sort_radix_f32(hpcs_ctx *ctx, float *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] {
        dev_sort<uint32_t, false, true>(ctx, reinterpret_cast<uint32_t *>(d_keys), nullptr, n, pick_stream(ctx, stream));
    });
}
#endif
#if HPCS_IN_PART(2)
HPCS_API int32_t hpcs_dev_sort_radix_f64(hpcs_ctx *ctx, double *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] {
        dev_sort<uint64_t, false, true>(ctx, reinterpret_cast<uint64_t *>(d_keys), nullptr, n, pick_stream(ctx, stream));
    });
}
#endif
#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_dev_sort_radix_pairs_u64_u32(hpcs_ctx *ctx, uint64_t *d_keys, uint32_t *d_vals, int64_t n,
                                                   void *stream)
{
    return guarded(ctx, [&] { dev_sort<uint64_t, true, false>(ctx, d_keys, d_vals, n, pick_stream(ctx, stream)); });
}
#endif
#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_dev_sort_radix_pairs_u32_u32(hpcs_ctx *ctx, uint32_t *d_keys, uint32_t *d_vals, int64_t n,
                                                   void *stream)
{
    return guarded(ctx, [&] { dev_sort<uint32_t, true, false>(ctx, d_keys, d_vals, n, pick_stream(ctx, stream)); });
}
#endif
#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_dev_sort_radix_records(hpcs_ctx *ctx, void *d_recs, int64_t n, int32_t rec_size,
                                             int32_t key_offset, int32_t key_bits, int32_t key_kind, void *stream)
{
    return guarded(ctx, [&] {
        dev_records(ctx, d_recs, n, rec_size, key_offset, key_bits, key_kind, pick_stream(ctx, stream));
    });
}
#endif
#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_byte_components_u32(hpcs_ctx *ctx, const uint32_t *d_keys, int64_t n,
                                                        uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] { dev_histogram<uint32_t, false>(ctx, d_keys, n, d_counts, pick_stream(ctx, stream)); });
}
#endif
#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_byte_components_u64(hpcs_ctx *ctx, const uint64_t *d_keys, int64_t n,
                                                        uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] { dev_histogram<uint64_t, false>(ctx, d_keys, n, d_counts, pick_stream(ctx, stream)); });
}
#endif

// ---------------------------------------------------------------------------
// Radix top-k
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_select_topk_u32(hpcs_ctx *ctx, uint32_t *a, int64_t start, int64_t length, int64_t k)
{
    return guarded(ctx, [&] { host_select_topk<uint32_t, false>(ctx, a, start, length, k); });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_topk_u32(hpcs_ctx *ctx, const uint32_t *keys, int64_t n, int64_t k, uint32_t *out)
{
    return guarded(ctx, [&] { host_topk<uint32_t, false>(ctx, keys, n, k, out); });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_topk_u64(hpcs_ctx *ctx, const uint64_t *keys, int64_t n, int64_t k, uint64_t *out)
{
    return guarded(ctx, [&] { host_topk<uint64_t, false>(ctx, keys, n, k, out); });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_topk_f32(hpcs_ctx *ctx, const float *keys, int64_t n, int64_t k, float *out)
{
    return guarded(ctx, [&] {
        host_topk<uint32_t, true>(ctx, reinterpret_cast<const uint32_t *>(keys), n, k, reinterpret_cast<uint32_t *>(out));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_topk_f64(hpcs_ctx *ctx, const double *keys, int64_t n, int64_t k, double *out)
{
    return guarded(ctx, [&] {
        host_topk<uint64_t, true>(ctx, reinterpret_cast<const uint64_t *>(keys), n, k, reinterpret_cast<uint64_t *>(out));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_topk_u32(hpcs_ctx *ctx, const uint32_t *d_keys, int64_t n, int64_t k, uint32_t *d_out,
                                   void *stream)
{
    return guarded(ctx, [&] { run_topk<uint32_t, false>(ctx, d_keys, n, k, d_out, nullptr, true, pick_stream(ctx, stream)); });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_topk_u64(hpcs_ctx *ctx, const uint64_t *d_keys, int64_t n, int64_t k, uint64_t *d_out,
                                   void *stream)
{
    return guarded(ctx, [&] { run_topk<uint64_t, false>(ctx, d_keys, n, k, d_out, nullptr, true, pick_stream(ctx, stream)); });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_topk_f32(hpcs_ctx *ctx, const float *d_keys, int64_t n, int64_t k, float *d_out, void *stream)
{
    return guarded(ctx, [&] {
        run_topk<uint32_t, true>(ctx, reinterpret_cast<const uint32_t *>(d_keys), n, k,
                                 reinterpret_cast<uint32_t *>(d_out), nullptr, true, pick_stream(ctx, stream));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_topk_f64(hpcs_ctx *ctx, const double *d_keys, int64_t n, int64_t k, double *d_out,
                                   void *stream)
{
    return guarded(ctx, [&] {
        run_topk<uint64_t, true>(ctx, reinterpret_cast<const uint64_t *>(d_keys), n, k,
                                 reinterpret_cast<uint64_t *>(d_out), nullptr, true, pick_stream(ctx, stream));
    });
}
#endif

// ---------------------------------------------------------------------------
// Signed keys: int[] / long[] (RadixSortLsd.cs:267-357, 444-532;
// RadixSortLsdParallel.cs:446-668; Histogram.cs:992-1052): the sign bit is
// flipped on the way in and out, so only the top digit differs from the
// unsigned sort, as in the reference (its top digit ^ 128).
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(6)
HPCS_API int32_t hpcs_sort_radix_i32(hpcs_ctx *ctx, int32_t *keys, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint32_t, false, KIND_INT>(ctx, reinterpret_cast<uint32_t *>(keys), nullptr, n); });
}
#endif

#if HPCS_IN_PART(2)
HPCS_API int32_t hpcs_sort_radix_i64(hpcs_ctx *ctx, int64_t *keys, int64_t n)
{
    return guarded(ctx, [&] { host_sort<uint64_t, false, KIND_INT>(ctx, reinterpret_cast<uint64_t *>(keys), nullptr, n); });
}
#endif

#if HPCS_IN_PART(6)
HPCS_API int32_t hpcs_dev_sort_radix_i32(hpcs_ctx *ctx, int32_t *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] {
        dev_sort<uint32_t, false, KIND_INT>(ctx, reinterpret_cast<uint32_t *>(d_keys), nullptr, n, pick_stream(ctx, stream));
    });
}
#endif

#if HPCS_IN_PART(2)
HPCS_API int32_t hpcs_dev_sort_radix_i64(hpcs_ctx *ctx, int64_t *d_keys, int64_t n, void *stream)
{
    return guarded(ctx, [&] {
        dev_sort<uint64_t, false, KIND_INT>(ctx, reinterpret_cast<uint64_t *>(d_keys), nullptr, n, pick_stream(ctx, stream));
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_byte_components_i32(hpcs_ctx *ctx, const int32_t *a, int64_t l, int64_t r,
                                                    int32_t *counts)
{
    return guarded(ctx, [&] {
        host_histogram<uint32_t, KIND_INT>(ctx, reinterpret_cast<const uint32_t *>(a), l, r, counts);
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_byte_components_i64(hpcs_ctx *ctx, const int64_t *a, int64_t l, int64_t r,
                                                    int32_t *counts)
{
    return guarded(ctx, [&] {
        host_histogram<uint64_t, KIND_INT>(ctx, reinterpret_cast<const uint64_t *>(a), l, r, counts);
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_byte_components_i32(hpcs_ctx *ctx, const int32_t *d_keys, int64_t n,
                                                        uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] {
        dev_histogram<uint32_t, KIND_INT>(ctx, reinterpret_cast<const uint32_t *>(d_keys), n, d_counts,
                                          pick_stream(ctx, stream));
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_byte_components_i64(hpcs_ctx *ctx, const int64_t *d_keys, int64_t n,
                                                        uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] {
        dev_histogram<uint64_t, KIND_INT>(ctx, reinterpret_cast<const uint64_t *>(d_keys), n, d_counts,
                                          pick_stream(ctx, stream));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_topk_i32(hpcs_ctx *ctx, const int32_t *keys, int64_t n, int64_t k, int32_t *out)
{
    return guarded(ctx, [&] {
        host_topk<uint32_t, KIND_INT>(ctx, reinterpret_cast<const uint32_t *>(keys), n, k, reinterpret_cast<uint32_t *>(out));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_topk_i64(hpcs_ctx *ctx, const int64_t *keys, int64_t n, int64_t k, int64_t *out)
{
    return guarded(ctx, [&] {
        host_topk<uint64_t, KIND_INT>(ctx, reinterpret_cast<const uint64_t *>(keys), n, k, reinterpret_cast<uint64_t *>(out));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_topk_i32(hpcs_ctx *ctx, const int32_t *d_keys, int64_t n, int64_t k, int32_t *d_out,
                                   void *stream)
{
    return guarded(ctx, [&] {
        run_topk<uint32_t, KIND_INT>(ctx, reinterpret_cast<const uint32_t *>(d_keys), n, k,
                                     reinterpret_cast<uint32_t *>(d_out), nullptr, true, pick_stream(ctx, stream));
    });
}
#endif

#if HPCS_IN_PART(4)
HPCS_API int32_t hpcs_dev_topk_i64(hpcs_ctx *ctx, const int64_t *d_keys, int64_t n, int64_t k, int64_t *d_out,
                                   void *stream)
{
    return guarded(ctx, [&] {
        run_topk<uint64_t, KIND_INT>(ctx, reinterpret_cast<const uint64_t *>(d_keys), n, k,
                                     reinterpret_cast<uint64_t *>(d_out), nullptr, true, pick_stream(ctx, stream));
    });
}
#endif

// ---------------------------------------------------------------------------
// Key-extractor histograms: HistogramByteComponents<T>(T[], l, r,
// Func<T, UInt32 / UInt64>) (Histogram.cs:760-844) with the key at a fixed
// offset inside each record.
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_byte_components_records(hpcs_ctx *ctx, const void *recs, int64_t l, int64_t r,
                                                        int32_t rec_size, int32_t key_offset, int32_t key_bits,
                                                        int32_t *counts)
{
    return guarded(ctx, [&] {
        if (!recs) fail(HPCS_E_INVALID_ARG, "inArray is null");
        if (!counts) fail(HPCS_E_INVALID_ARG, "counts is null");
        if (l < 0) fail(HPCS_E_INVALID_ARG, "l must be non-negative");
        const int64_t n = r >= l ? r - l + 1 : 0;
        check_records_args(n, rec_size, key_offset, key_bits, HPCS_KEY_UINT);
        if (n >= (int64_t(1) << 31)) fail(HPCS_E_UNSUPPORTED, "int32 counts limit the range to 2^31 - 1 keys");
        const int P = key_bits / 8;
        const size_t hoff = align_up(size_t(n) * size_t(rec_size));
        char *io = static_cast<char *>(ensure(ctx->io, ctx->io_cap, hoff + P * RADIX * sizeof(uint32_t)));
        uint32_t *d_counts = reinterpret_cast<uint32_t *>(io + hoff);
        hipStream_t st = ctx->stream;
        if (n > 0)
            hip_check(hipMemcpyAsync(io, static_cast<const unsigned char *>(recs) + size_t(l) * size_t(rec_size),
                                     size_t(n) * size_t(rec_size), hipMemcpyHostToDevice, st),
                      "H2D");
        dev_histogram_records(ctx, io, n, rec_size, key_offset, key_bits, d_counts, st);
        uint32_t tmp[8 * RADIX];
        hip_check(hipMemcpyAsync(tmp, d_counts, P * RADIX * sizeof(uint32_t), hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        for (int i = 0; i < P * RADIX; ++i) counts[i] = int32_t(tmp[i]);
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_byte_components_records(hpcs_ctx *ctx, const void *d_recs, int64_t n,
                                                            int32_t rec_size, int32_t key_offset, int32_t key_bits,
                                                            uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] {
        dev_histogram_records(ctx, d_recs, n, rec_size, key_offset, key_bits, d_counts, pick_stream(ctx, stream));
    });
}
#endif

// ---------------------------------------------------------------------------
// Tuning tables
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_ctx_load_tuning(hpcs_ctx *ctx, const char *path)
{
    return guarded(ctx, [&] {
        graph_cache_clear(ctx); // cached launch sequences follow the table
        if (!path || !*path) {
            ctx->tune.clear();
            return;
        }
        std::string text, err;
        std::vector<TuneEntry> rows;
        if (!read_file(path, text)) fail(HPCS_E_INVALID_ARG, std::string("cannot read ") + path);
        if (!parse_tuning(text, rows, err)) fail(HPCS_E_INVALID_ARG, err);
        ctx->tune.swap(rows);
    });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_tuning_check(const char *path, int32_t *rows)
{
    if (!path) {
        g_last_error = "path is null";
        return HPCS_E_INVALID_ARG;
    }
    std::string text, err;
    std::vector<TuneEntry> out;
    if (!read_file(path, text)) {
        g_last_error = std::string("cannot read ") + path;
        return HPCS_E_INVALID_ARG;
    }
    if (!parse_tuning(text, out, err)) {
        g_last_error = err;
        return HPCS_E_INVALID_ARG;
    }
    if (rows) *rows = int32_t(out.size());
    return HPCS_OK;
}
#endif

// ---------------------------------------------------------------------------
// N-bit histograms: Histogram11bitComponents(uint[], l, r) (Histogram.cs:
// 298-331, bits = 11) and HistogramNbitComponents(uint[] / ulong[], l, r,
// bits) (:381-427)
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_nbit_components_u32(hpcs_ctx *ctx, const uint32_t *a, int64_t l, int64_t r,
                                                    int32_t bits, int32_t *counts)
{
    return guarded(ctx, [&] { host_histogram_nbit<uint32_t>(ctx, a, l, r, bits, counts); });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_histogram_nbit_components_u64(hpcs_ctx *ctx, const uint64_t *a, int64_t l, int64_t r,
                                                    int32_t bits, int32_t *counts)
{
    return guarded(ctx, [&] { host_histogram_nbit<uint64_t>(ctx, a, l, r, bits, counts); });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_nbit_components_u32(hpcs_ctx *ctx, const uint32_t *d_keys, int64_t n,
                                                        int32_t bits, uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] { dev_histogram_nbit<uint32_t>(ctx, d_keys, n, bits, d_counts, pick_stream(ctx, stream)); });
}
#endif

#if HPCS_IN_PART(0)
HPCS_API int32_t hpcs_dev_histogram_nbit_components_u64(hpcs_ctx *ctx, const uint64_t *d_keys, int64_t n,
                                                        int32_t bits, uint32_t *d_counts, void *stream)
{
    return guarded(ctx, [&] { dev_histogram_nbit<uint64_t>(ctx, d_keys, n, bits, d_counts, pick_stream(ctx, stream)); });
}
#endif

// ---------------------------------------------------------------------------
// Items by a separate key array: SortRadix<T>(this T[] inputArray, UInt32[] /
// UInt64[] inKeys) (RadixSortLsd.cs:1183-1299) for items of any fixed size
// ---------------------------------------------------------------------------

#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_sort_radix_keys_items(hpcs_ctx *ctx, void *keys, int32_t key_bits, void *items, int32_t item_size,
                                            int64_t n)
{
    return guarded(ctx, [&] {
        check_keys_items_args(n, key_bits, item_size);
        if (n <= 1) return;
        if (!keys || !items) fail(HPCS_E_INVALID_ARG, "keys or items is null");
        const size_t kb = size_t(n) * size_t(key_bits / 8);
        const size_t ioff = align_up(kb);
        char *io = static_cast<char *>(ensure(ctx->io, ctx->io_cap, ioff + size_t(n) * size_t(item_size)));
        hipStream_t st = ctx->stream;
        hip_check(hipMemcpyAsync(io, keys, kb, hipMemcpyHostToDevice, st), "H2D");
        hip_check(hipMemcpyAsync(io + ioff, items, size_t(n) * size_t(item_size), hipMemcpyHostToDevice, st), "H2D");
        dev_keys_items(ctx, io, key_bits, io + ioff, item_size, n, st);
        hip_check(hipMemcpyAsync(keys, io, kb, hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipMemcpyAsync(items, io + ioff, size_t(n) * size_t(item_size), hipMemcpyDeviceToHost, st), "D2H");
        hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    });
}
#endif

#if HPCS_IN_PART(3)
HPCS_API int32_t hpcs_dev_sort_radix_keys_items(hpcs_ctx *ctx, void *d_keys, int32_t key_bits, void *d_items,
                                                int32_t item_size, int64_t n, void *stream)
{
    return guarded(ctx, [&] { dev_keys_items(ctx, d_keys, key_bits, d_items, item_size, n, pick_stream(ctx, stream)); });
}
#endif
