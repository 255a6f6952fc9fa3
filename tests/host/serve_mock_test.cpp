// Host test of the serving core (hpcsharp_amd/csrc/hpcs_serve_core.hpp)
// with a mock executor: each slot's "device" is a thread that runs the
// enqueued work in order, taking a time proportional to its keys, and
// completes its tokens.  Checked: every result and its timestamps, the
// budget and the arena bounds, coalesced batches (consecutive, within their
// limits, only small 4-byte-key sorts), rejection of oversize requests,
// injected allocation failures (at set-up and per request), and the
// completion deadline (a request that never completes, and one that is
// merely slower than a tiny deadline).  Built with -fsanitize=thread (and
// separately with address/undefined) and run by tests/test_serving_host.py.
#include "../../hpcsharp_amd/csrc/hpcs_serve_core.hpp"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <random>

using namespace hpcs_serve;

namespace {

constexpr uint64_t HANG_SEED = 0x4A4E47ull; // the mock device never completes this request
constexpr uint64_t OOM_SEED = 0x4F4F4Dull;  // the mock's sort call fails with HPCS_E_OOM

uint64_t mix(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t expected_sum(const hpcs_request &r) { return mix(r.seed ^ (uint64_t(r.n) << 1) ^ uint64_t(r.op)); }

int g_failures = 0;
#define CHECK(cond)                                                                                                    \
    do {                                                                                                               \
        if (!(cond)) {                                                                                                 \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);                                               \
            ++g_failures;                                                                                              \
        }                                                                                                              \
    } while (0)

struct MockExec {
    struct Work {
        const hpcs_request *reqs;
        int token;
        std::vector<int64_t> idx;
        std::vector<int32_t> status;
        double seconds;
        bool hang;
    };
    struct Slot {
        std::mutex m;
        std::condition_variable cv;
        std::deque<Work> q;
        bool stop = false, busy = false;
        std::atomic<bool> done[SLOT_DEPTH];
        std::thread dev;
    };

    // options
    bool fail_setup = false;
    double ns_per_key = 2.0, min_seconds = 20e-6;

    // state
    std::vector<uint64_t> sums;
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<char> arena;
    int64_t seg_keys = 0;
    int seg_requests = 0;

    // record (guarded by rec_m)
    std::mutex rec_m;
    std::vector<std::vector<int64_t>> batches;
    int bad_blocks = 0;
    bool leaked = false, released = false;

    int64_t slot_scratch(int32_t dtype, int64_t n) const
    {
        return n > 0 ? int64_t(2 * size_t(n) * key_size(dtype)) + (int64_t(1) << 20) : 0;
    }

    void setup(int nslots, const int64_t nmax[2], int64_t batch_keys, int batch_requests, size_t arena_bytes,
               int64_t nreq)
    {
        (void)nmax;
        if (fail_setup) throw Fail{HPCS_E_OOM, "injected: arena allocation failed"};
        seg_keys = batch_keys;
        seg_requests = batch_requests;
        arena.assign(std::max<size_t>(arena_bytes, 1), 0);
        sums.assign(size_t(std::max<int64_t>(nreq, 1)), 0);
        for (int s = 0; s < nslots; ++s) {
            slots.emplace_back(new Slot());
            Slot &sl = *slots.back();
            for (auto &d : sl.done) d.store(false);
            sl.dev = std::thread([this, &sl] { device(sl); });
        }
    }

    char *arena_base() { return arena.data(); }

    void device(Slot &sl)
    {
        for (;;) {
            Work w;
            {
                std::unique_lock<std::mutex> lk(sl.m);
                sl.cv.wait(lk, [&] { return sl.stop || !sl.q.empty(); });
                if (sl.q.empty()) return; // stopped
                if (sl.q.front().hang) {   // this work never completes: wait for the stop
                    sl.busy = true;
                    sl.cv.wait(lk, [&] { return sl.stop; });
                    return;
                }
                w = sl.q.front();
                sl.busy = true;
            }
            std::this_thread::sleep_for(std::chrono::duration<double>(w.seconds));
            for (size_t k = 0; k < w.idx.size(); ++k)
                sums[size_t(w.idx[k])] = w.status[k] == HPCS_OK ? expected_sum(w.reqs[w.idx[k]]) : 0;
            {
                std::lock_guard<std::mutex> lk(sl.m);
                sl.q.pop_front();
                sl.busy = false;
            }
            sl.done[w.token].store(true, std::memory_order_release);
            sl.cv.notify_all();
        }
    }

    void submit(int s, int token, const hpcs_request *r, const int64_t *idx, int count, char *block, int32_t *status)
    {
        Slot &sl = *slots[size_t(s)];
        const size_t bytes = batch_bytes(r, idx, count);
        const bool inside = block >= arena.data() && block + bytes <= arena.data() + arena.size();
        int64_t keys = 0;
        Work w{r, token, std::vector<int64_t>(idx, idx + count), std::vector<int32_t>(size_t(count), HPCS_OK), 0.0,
               false};
        for (int k = 0; k < count; ++k) {
            keys += r[idx[k]].n;
            if (r[idx[k]].seed == OOM_SEED) status[k] = w.status[size_t(k)] = HPCS_E_OOM;
            if (r[idx[k]].seed == HANG_SEED) w.hang = true;
        }
        w.seconds = std::max(min_seconds, double(keys) * ns_per_key * 1e-9);
        {
            std::lock_guard<std::mutex> lk(rec_m);
            batches.push_back(w.idx);
            if (!inside) ++bad_blocks;
        }
        sl.done[token].store(false, std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> lk(sl.m);
            sl.q.push_back(std::move(w));
        }
        sl.cv.notify_all();
    }

    bool poll(int s, int token) { return slots[size_t(s)]->done[token].load(std::memory_order_acquire); }

    // Polls rather than a timed condition wait: this toolchain's thread
    // sanitizer does not see the mutex released inside a timed wait.
    bool drain(int s, double seconds)
    {
        Slot &sl = *slots[size_t(s)];
        const auto end = Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(seconds));
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(sl.m);
                if (sl.q.empty() && !sl.busy) return true;
            }
            if (Clock::now() > end) return false;
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    }

    void checksums(uint64_t *out, int64_t nreq)
    {
        for (int64_t i = 0; i < nreq; ++i) out[i] = sums[size_t(i)];
    }

    void release(bool leak)
    {
        for (auto &p : slots) {
            {
                std::lock_guard<std::mutex> lk(p->m);
                p->stop = true;
            }
            p->cv.notify_all();
            p->dev.join();
        }
        slots.clear();
        std::lock_guard<std::mutex> lk(rec_m);
        leaked = leak;
        released = true;
    }
};

hpcs_serve_config config(int slots, int64_t budget)
{
    hpcs_serve_config c{};
    c.slots = slots;
    c.mode = HPCS_SERVE_REALTIME;
    c.hbm_budget = budget;
    return c;
}

// A mixed trace: sizes log-uniform in [1, nmax], sorts and top-k, all dtypes.
std::vector<hpcs_request> trace(int64_t nreq, double span_s, int64_t nmax, uint64_t seed, double sort_frac = 0.7)
{
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<hpcs_request> t(static_cast<size_t>(nreq));
    std::vector<double> at(static_cast<size_t>(nreq));
    for (auto &a : at) a = u(g) * span_s;
    std::sort(at.begin(), at.end());
    for (int64_t i = 0; i < nreq; ++i) {
        hpcs_request &r = t[size_t(i)];
        r.op = u(g) < sort_frac ? HPCS_OP_SORT : HPCS_OP_TOPK;
        r.dtype = int32_t(g() % 4);
        r.n = std::max<int64_t>(1, int64_t(std::exp(u(g) * std::log(double(nmax)))));
        r.k = r.op == HPCS_OP_TOPK ? std::max<int64_t>(1, r.n / 3) : 0;
        r.seed = g() | 0x1000000ull; // never a magic seed
        r.arrival_s = at[size_t(i)];
    }
    return t;
}

struct Outcome {
    int32_t rc;
    std::string err;
    std::vector<hpcs_request_result> res;
    hpcs_serve_stats stats;
    double seconds;
};

Outcome run(MockExec &ex, const hpcs_serve_config &cfg, const std::vector<hpcs_request> &t)
{
    Outcome o{};
    o.res.assign(t.size(), hpcs_request_result{});
    std::memset(&o.stats, 0, sizeof(o.stats));
    const auto t0 = Clock::now();
    try {
        const hpcs_serve_config c = resolve_config(cfg);
        validate(t.data(), int64_t(t.size()));
        Server<MockExec> S(ex, c, t.data(), int64_t(t.size()), o.res.data());
        o.rc = S.run(&o.stats, &o.err);
    } catch (const Fail &e) {
        o.rc = e.code;
        o.err = e.msg;
    }
    o.seconds = std::chrono::duration<double>(Clock::now() - t0).count();
    return o;
}

// Every completed request: its checksum, ordered timestamps; with
// realtime arrivals, admission after arrival.
void check_results(const std::vector<hpcs_request> &t, const Outcome &o, bool realtime)
{
    for (size_t i = 0; i < t.size(); ++i) {
        const hpcs_request_result &r = o.res[i];
        if (r.status != HPCS_OK) continue;
        CHECK(r.checksum == expected_sum(t[i]));
        CHECK(r.enqueued_s >= r.admit_s);
        CHECK(r.done_s >= r.enqueued_s);
        if (realtime) CHECK(r.admit_s >= t[i].arrival_s - 1e-9);
    }
    // first come, first served: admissions in request order
    for (size_t i = 1; i < t.size(); ++i) CHECK(o.res[i].admit_s >= o.res[i - 1].admit_s);
}

void test_mixed_realtime()
{
    MockExec ex;
    const auto t = trace(300, 0.05, 100000, 1);
    const Outcome o = run(ex, config(4, int64_t(1) << 30), t);
    CHECK(o.rc == HPCS_OK);
    check_results(t, o, true);
    int used[4] = {0, 0, 0, 0};
    for (size_t i = 0; i < t.size(); ++i) {
        CHECK(o.res[i].status == HPCS_OK);
        if (o.res[i].slot >= 0 && o.res[i].slot < 4) used[o.res[i].slot] = 1;
    }
    CHECK(used[0] + used[1] + used[2] + used[3] >= 2);
    CHECK(o.stats.batches == 300 && o.stats.coalesced_requests == 0);
    CHECK(o.stats.peak_bytes <= o.stats.request_budget_bytes);
    CHECK(ex.bad_blocks == 0 && ex.released && !ex.leaked);
}

void test_coalescing()
{
    MockExec ex;
    ex.min_seconds = 200e-6;
    auto t = trace(600, 0.0, 4096, 2, 1.0); // every request arrives at once
    for (size_t i = 0; i < t.size(); ++i)
        if (i % 7 == 3) t[i].dtype = HPCS_DTYPE_U64;           // not coalescable: ends a run
        else if (i % 11 == 5) t[i].op = HPCS_OP_TOPK, t[i].k = 1;
    t[100].n = 5000;                                           // above the size limit
    hpcs_serve_config c = config(2, int64_t(1) << 30);
    c.coalesce_max_keys = 4096;
    c.coalesce_max_requests = 16;
    c.coalesce_max_batch_keys = 20000;
    const Outcome o = run(ex, c, t);
    CHECK(o.rc == HPCS_OK);
    check_results(t, o, true);
    int64_t in_batches = 0, multi = 0;
    for (const auto &b : ex.batches) {
        if (b.size() < 2) continue;
        ++multi;
        in_batches += int64_t(b.size());
        CHECK(b.size() <= 16);
        int64_t keys = 0;
        for (size_t k = 0; k < b.size(); ++k) {
            const hpcs_request &r = t[size_t(b[k])];
            keys += r.n;
            CHECK(r.op == HPCS_OP_SORT && (r.dtype == HPCS_DTYPE_U32 || r.dtype == HPCS_DTYPE_F32) && r.n <= 4096);
            if (k) CHECK(b[k] == b[k - 1] + 1); // consecutive in the queue
        }
        CHECK(keys <= 20000);
    }
    CHECK(multi > 10);
    CHECK(o.stats.coalesced_requests == in_batches);
    CHECK(o.stats.batches == int64_t(ex.batches.size()));
    for (size_t i = 0; i < t.size(); ++i) CHECK(o.res[i].status == HPCS_OK);
    CHECK(ex.bad_blocks == 0);
}

void test_budget()
{
    {
        MockExec ex; // no request fits beside its scratch: each is rejected, the run succeeds
        const auto t = trace(20, 0.0, 1 << 20, 3);
        const Outcome o = run(ex, config(4, int64_t(1) << 20), t);
        CHECK(o.rc == HPCS_OK);
        for (size_t i = 0; i < t.size(); ++i) CHECK(o.res[i].status == HPCS_E_BUDGET);
        CHECK(o.stats.keys == 0);
    }
    {
        MockExec ex; // one request too large for the budget beside the others' scratch: rejected alone
        auto t = trace(50, 0.01, 1000, 4);
        t[17].n = 40000000;
        t[17].op = HPCS_OP_SORT;
        t[17].dtype = HPCS_DTYPE_U64;
        const Outcome o = run(ex, config(2, int64_t(400) << 20), t);
        CHECK(o.rc == HPCS_OK);
        check_results(t, o, true);
        for (size_t i = 0; i < t.size(); ++i) CHECK(o.res[i].status == (i == 17 ? HPCS_E_BUDGET : HPCS_OK));
        CHECK(o.stats.peak_bytes <= o.stats.request_budget_bytes);
    }
    {
        MockExec ex; // a tight arena: few requests fit at once, all complete
        ex.ns_per_key = 50.0;
        const auto t = trace(200, 0.0, 100000, 5);
        const int64_t scratch = ex.slot_scratch(HPCS_DTYPE_U64, 100000);
        const Outcome o = run(ex, config(3, scratch * 3 + (int64_t(2) << 20)), t);
        CHECK(o.rc == HPCS_OK);
        check_results(t, o, true);
        for (size_t i = 0; i < t.size(); ++i) CHECK(o.res[i].status == HPCS_OK);
        CHECK(o.stats.peak_bytes <= o.stats.request_budget_bytes);
        CHECK(ex.bad_blocks == 0);
    }
}

void test_injected_allocation_failures()
{
    {
        MockExec ex;
        ex.fail_setup = true;
        const auto t = trace(10, 0.0, 1000, 6);
        const Outcome o = run(ex, config(2, int64_t(1) << 30), t);
        CHECK(o.rc == HPCS_E_OOM);
        CHECK(o.err.find("injected") != std::string::npos);
        CHECK(ex.released && !ex.leaked);
    }
    {
        MockExec ex; // the sort call of two requests fails: they report it, the rest complete
        auto t = trace(100, 0.01, 10000, 7);
        t[10].seed = OOM_SEED;
        t[60].seed = OOM_SEED;
        const Outcome o = run(ex, config(4, int64_t(1) << 30), t);
        CHECK(o.rc == HPCS_OK);
        check_results(t, o, true);
        for (size_t i = 0; i < t.size(); ++i)
            CHECK(o.res[i].status == ((i == 10 || i == 60) ? HPCS_E_OOM : HPCS_OK));
        CHECK(o.res[10].checksum == 0 && o.res[60].checksum == 0);
    }
}

void test_deadlines()
{
    {
        MockExec ex; // request 25 never completes: the run fails naming it and keeps the memory
        auto t = trace(60, 0.01, 10000, 8);
        t[25].seed = HANG_SEED;
        hpcs_serve_config c = config(2, int64_t(1) << 30);
        c.request_timeout_s = 0.2;
        const Outcome o = run(ex, c, t);
        CHECK(o.rc == HPCS_E_HIP);
        CHECK(o.err.find("request 25 ") != std::string::npos);
        CHECK(o.err.find("did not complete") != std::string::npos);
        CHECK(o.seconds < 5.0);
        CHECK(ex.released && ex.leaked);
    }
    {
        MockExec ex; // work slower than a tiny deadline: the run fails, the drain ends, nothing leaks
        ex.min_seconds = 2e-3;
        const auto t = trace(40, 0.0, 1000, 9);
        hpcs_serve_config c = config(2, int64_t(1) << 30);
        c.request_timeout_s = 1e-6;
        const Outcome o = run(ex, c, t);
        CHECK(o.rc == HPCS_E_HIP);
        CHECK(o.err.find("did not complete within") != std::string::npos);
        CHECK(ex.released && !ex.leaked);
    }
}

void test_bad_configs()
{
    {
        MockExec ex; // coalesced batches that cannot fit the budget left after the slots' scratch
        const auto t = trace(4, 0.0, 100, 12);
        // the slots' scratch (sized for 2^22-key batches) fits, 8 MiB are left: less than one batch
        hpcs_serve_config c = config(2, 2 * ex.slot_scratch(HPCS_DTYPE_U32, int64_t(1) << 22) + (int64_t(8) << 20));
        c.coalesce_max_keys = 1024;
        const Outcome o = run(ex, c, t);
        CHECK(o.rc == HPCS_E_BUDGET);
        CHECK(o.err.find("coalesced") != std::string::npos);
    }
    const auto t = trace(4, 0.0, 100, 10);
    hpcs_serve_config c = config(2, int64_t(1) << 30);
    c.coalesce_max_requests = 1;
    MockExec a;
    CHECK(run(a, c, t).rc == HPCS_E_INVALID_ARG);
    c = config(2, int64_t(1) << 30);
    c.request_timeout_s = -1.0;
    MockExec b;
    CHECK(run(b, c, t).rc == HPCS_E_INVALID_ARG);
    c = config(0, int64_t(1) << 30);
    MockExec d;
    CHECK(run(d, c, t).rc == HPCS_E_INVALID_ARG);
}

void test_stress()
{
    // many slots, a budget that keeps admission contended, coalescing on
    MockExec ex;
    ex.ns_per_key = 5.0;
    auto t = trace(3000, 0.2, 200000, 11);
    hpcs_serve_config c = config(8, int64_t(0));
    c.hbm_budget = ex.slot_scratch(HPCS_DTYPE_U64, 200000) * 8 + (int64_t(8) << 20);
    c.coalesce_max_keys = 2048;
    c.coalesce_max_batch_keys = 16384;
    const Outcome o = run(ex, c, t);
    CHECK(o.rc == HPCS_OK);
    check_results(t, o, true);
    for (size_t i = 0; i < t.size(); ++i) CHECK(o.res[i].status == HPCS_OK);
    CHECK(o.stats.peak_bytes <= o.stats.request_budget_bytes);
    CHECK(ex.bad_blocks == 0);
}

} // namespace

int main()
{
    test_mixed_realtime();
    test_coalescing();
    test_budget();
    test_injected_allocation_failures();
    test_deadlines();
    test_bad_configs();
    test_stress();
    if (g_failures) {
        std::printf("%d check(s) failed\n", g_failures);
        return 1;
    }
    std::printf("serve core: all checks passed\n");
    return 0;
}
