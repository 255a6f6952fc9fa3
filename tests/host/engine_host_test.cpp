// Host-side checks of the engine's plain C++ parts (built with g++ and the
// address / undefined-behaviour sanitizers by tests/test_host_engine.py):
// the serving arena's first-fit allocator against a brute-force model, and
// the tuning-table parser.
#include "../../hpcsharp_amd/csrc/hpcs_arena.hpp"
#include "../../hpcsharp_amd/csrc/hpcs_tuning.hpp"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#define CHECK(c)                                                                 \
    do {                                                                         \
        if (!(c)) {                                                              \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                            \
        }                                                                        \
    } while (0)

static int arena_test()
{
    const size_t unit = 256, units = 4096;
    hpcs::Arena a;
    a.reset(reinterpret_cast<char *>(0x1000), units * unit);
    std::vector<int> owner(units, -1); // model: which live block owns each unit
    struct Blk {
        size_t off, bytes;
    };
    std::vector<Blk> live;
    std::mt19937_64 rng(12345);
    for (int step = 0; step < 200000; ++step) {
        const bool do_alloc = live.empty() || (rng() % 100) < 55;
        if (do_alloc) {
            const size_t n = 1 + rng() % 64;
            const size_t off = a.alloc(n * unit);
            // first fit: the lowest run of n free units
            size_t want = SIZE_MAX;
            for (size_t u = 0, run = 0; u < units; ++u) {
                run = owner[u] < 0 ? run + 1 : 0;
                if (run == n) {
                    want = (u + 1 - n) * unit;
                    break;
                }
            }
            CHECK(off == want);
            if (off != SIZE_MAX) {
                for (size_t u = off / unit; u < off / unit + n; ++u) {
                    CHECK(owner[u] < 0);
                    owner[u] = int(live.size());
                }
                live.push_back({off, n * unit});
                CHECK(a.at(off) == reinterpret_cast<char *>(0x1000) + off);
            }
        } else {
            const size_t k = rng() % live.size();
            const Blk b = live[k];
            a.release(b.off, b.bytes);
            for (size_t u = b.off / unit; u < (b.off + b.bytes) / unit; ++u) owner[u] = -1;
            live[k] = live.back();
            live.pop_back();
            for (size_t j = 0; j < live.size(); ++j)
                for (size_t u = live[j].off / unit; u < (live[j].off + live[j].bytes) / unit; ++u) owner[u] = int(j);
        }
    }
    for (const Blk &b : live) a.release(b.off, b.bytes);
    CHECK(a.alloc(units * unit) == 0); // everything merged back into one block
    return 0;
}

static int tuning_test()
{
    std::vector<hpcs::TuneEntry> rows;
    std::string err;
    CHECK(hpcs::parse_tuning("# c\nu32 4096 0 1 1  # x\n\n  pairs64 100 1 0 0\r\nf64 268435456 1 0 1", rows, err));
    CHECK(rows.size() == 3);
    CHECK(rows[0].key_bytes == 4 && rows[0].vals == 0 && rows[0].max_n == 4096 && rows[0].shape == 1);
    CHECK(rows[1].key_bytes == 8 && rows[1].vals == 1 && rows[1].algo == 1);
    CHECK(rows[2].max_n == 268435456);
    CHECK(hpcs::parse_tuning("u64 1 0 4 1\npairs32 7 1 4 0", rows, err) && rows.size() == 2 && rows[1].shape == 4);
    CHECK(rows[0].bits == 8 && rows[1].bits == 8);
    CHECK(hpcs::parse_tuning("u32 65536 0 0 0 11  # eleven\nf64 9 1 0 1 8", rows, err) && rows.size() == 2);
    CHECK(rows[0].bits == 11 && rows[0].algo == 0 && rows[1].bits == 8 && rows[1].algo == 1);
    // rows of the shape table: 5 (1024x16) serves 4-byte keys, 6 (1024x8)
    // 8-byte keys and 4-byte keys with values, both with ballot ranking only
    CHECK(hpcs::parse_tuning("u32 1 1 5 0\npairs32 2 1 6 0\nf64 3 0 6 0", rows, err) && rows.size() == 3);
    CHECK(rows[0].shape == 5 && rows[1].shape == 6 && rows[2].shape == 6);
    // ranks 2 and 3: LDS-OR masks with every second / fourth item by ballot
    CHECK(hpcs::parse_tuning("u32 1 1 0 2\nu64 2 1 4 3", rows, err) && rows.size() == 2 && rows[1].rank == 3);
    const char *bad[] = {"u8 1 0 0 0", "u32 0 0 0 0", "u32 1 2 0 0", "u32 1 0 10 0", "u32 1 0 0 4", "u64 1 0 7 0", "u64 1 0 1 0",
                         "u32 1 1 5 1", "pairs32 1 1 5 0", "u32 1 1 6 0", "pairs64 1 1 6 0", "u32 1 0 -1 0",
                         "pairs32 1 0 3 1", "u32 1 0 0 1 9", "u32 1 0", "u32 -5 0 0 1",
                         "u32 1 1 0 0 11", "u32 1 0 4 0 11", "u32 1 0 0 1 11", "pairs64 1 0 0 0 11",
                         "u32 1 0 0 0 11 8", "u32 1 0 0 0 x",
                         "aaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa 1 0 0 1"};
    for (const char *b : bad) {
        CHECK(!hpcs::parse_tuning(b, rows, err));
        CHECK(err.rfind("line 1", 0) == 0);
    }
    CHECK(!hpcs::parse_tuning("u32 1 0 0 1\nbogus", rows, err) && err.rfind("line 2", 0) == 0);
    // batched-segment rows: segs <max_n> <small_max> <rank>
    CHECK(hpcs::parse_tuning("segs 1000 4096 1 # lds\nu32 9 1 0 3\n segs 268435456 16384 3\r", rows, err));
    CHECK(rows.size() == 3 && rows[0].key_bytes == 0 && rows[0].seg_small == 4096 && rows[0].rank == 1);
    CHECK(rows[1].key_bytes == 4 && rows[2].key_bytes == 0 && rows[2].max_n == 268435456 &&
          rows[2].seg_small == 16384 && rows[2].rank == 3 && rows[0].shape == 0 && rows[2].shape == 0);
    CHECK(hpcs::parse_tuning("segs 10 4096 2 9\nsegs 20 16384 0 3  # 512x32", rows, err) && rows.size() == 2);
    CHECK(rows[0].shape == 9 && rows[1].shape == 3 && rows[1].rank == 0);
    const char *bad_segs[] = {"segs 1 8192 1", "segs 0 4096 1", "segs 1 4096 4", "segs 1 4096 -1", "segs 1 4096",
                              "segs 1 4096 1 5", "segs 1 4096 1 6", "segs 1 4096 1 10", "segs 1 4096 1 0 0",
                              "segs 1 16384 x", "segs -3 4096 0", "segs 1 4096 1 x"};
    for (const char *b : bad_segs) {
        CHECK(!hpcs::parse_tuning(b, rows, err));
        CHECK(err.rfind("line 1", 0) == 0);
    }
    return 0;
}

int main()
{
    if (arena_test()) return 1;
    if (tuning_test()) return 1;
    std::printf("host engine checks passed\n");
    return 0;
}
