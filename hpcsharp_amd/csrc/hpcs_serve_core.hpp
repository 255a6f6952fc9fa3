// hpcs_serve_core.hpp -- the serving loop's host side: admission against the
// HBM budget, worker slots (a submitter and a reaper thread each), request
// coalescing and completion deadlines, written against an executor that owns
// the device.  Plain C++ (no HIP): hpcs_serve.hip instantiates it with the
// HIP executor, and the host test instantiates it with a mock executor under
// the thread sanitizer.
//
// The executor (template parameter Exec) provides:
//   int64_t slot_scratch(int32_t dtype, int64_t n)
//       device scratch one slot reserves for requests of up to n keys;
//   void setup(int slots, const int64_t nmax[2], int64_t batch_keys, int batch_requests,
//              size_t arena_bytes, int64_t nreq)
//       streams, contexts (scratch for nmax[0] 4-byte / nmax[1] 8-byte keys and
//       for segmented sorts of batch_keys keys in batch_requests segments),
//       the arena and the per-request checksums; throws Fail;
//   char *arena_base()
//   void submit(int slot, int token, const hpcs_request *reqs, const int64_t *idx, int count,
//               char *block, int32_t *status)
//       enqueue requests idx[0, count) (count > 1: a coalesced batch laid out
//       as BatchLayout in block) on the slot's stream, then mark completion
//       token `token` (0 <= token < SLOT_DEPTH); status[k] is the call's
//       status for idx[k]; throws Fail if the device rejects the launches;
//   bool poll(int slot, int token)   true once the token's work is complete; throws Fail
//   bool drain(int slot, double seconds)
//       wait for everything enqueued on the slot; false if the time ran out;
//   void checksums(uint64_t *out, int64_t nreq)   throws Fail
//   void release(bool leak)
//       free everything; leak: some work may still run, so device memory is
//       left allocated (the process keeps it) rather than freed under it.
#pragma once
#include "../../include/hpcsort.h"
#include "hpcs_arena.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace hpcs_serve {

struct Fail {
    int32_t code;
    std::string msg;
};

using Clock = std::chrono::steady_clock;

// Requests (or coalesced batches) a slot may have enqueued at once: the host
// prepares the next while the device runs the previous ones.
constexpr int SLOT_DEPTH = 4;
constexpr int COALESCE_MAX_REQUESTS = 1024;

inline size_t key_size(int32_t dtype) { return (dtype == HPCS_DTYPE_U32 || dtype == HPCS_DTYPE_F32) ? 4 : 8; }
inline size_t round_up(size_t x) { return (x + 255) / 256 * 256; }

// Device bytes a request holds in flight on its own: its payload and, for
// top-k, the output.
inline size_t request_bytes(const hpcs_request &r)
{
    const size_t ks = key_size(r.dtype);
    return round_up(size_t(r.n) * ks) + (r.op == HPCS_OP_TOPK ? round_up(size_t(r.k) * ks) : 0);
}

// A coalesced batch of count 4-byte-key sorts in one block: the payloads
// back to back (segment k is keys[begin[k], begin[k + 1])), then the
// segment starts (count + 1 int64), the seeds (count uint64), the request
// indices (count int64) and the float flags (count uint32).
struct BatchLayout {
    size_t off_begin, off_seed, off_ids, off_flt, total;
    int64_t keys;
};

inline BatchLayout batch_layout(const hpcs_request *reqs, const int64_t *idx, int count)
{
    BatchLayout L{};
    for (int k = 0; k < count; ++k) L.keys += reqs[idx[k]].n;
    L.off_begin = round_up(size_t(L.keys) * 4);
    L.off_seed = L.off_begin + round_up(size_t(count + 1) * 8);
    L.off_ids = L.off_seed + round_up(size_t(count) * 8);
    L.off_flt = L.off_ids + round_up(size_t(count) * 8);
    L.total = L.off_flt + round_up(size_t(count) * 4);
    return L;
}

inline size_t batch_bytes(const hpcs_request *reqs, const int64_t *idx, int count)
{
    return count == 1 ? request_bytes(reqs[idx[0]]) : batch_layout(reqs, idx, count).total;
}

// The configuration with its defaults filled in, checked.
inline hpcs_serve_config resolve_config(const hpcs_serve_config &in)
{
    hpcs_serve_config c = in;
    if (c.slots < 1 || c.slots > 64) throw Fail{HPCS_E_INVALID_ARG, "slots must lie in [1, 64]"};
    if (c.mode != HPCS_SERVE_REALTIME && c.mode != HPCS_SERVE_ALONE) throw Fail{HPCS_E_INVALID_ARG, "unknown mode"};
    if (c.hbm_budget <= 0) throw Fail{HPCS_E_INVALID_ARG, "budget must be positive"};
    if (!(c.request_timeout_s >= 0.0)) throw Fail{HPCS_E_INVALID_ARG, "request timeout must be non-negative"};
    if (c.coalesce_max_keys < 0 || c.coalesce_max_keys >= (int64_t(1) << 31))
        throw Fail{HPCS_E_INVALID_ARG, "coalesce_max_keys must lie in [0, 2^31)"};
    if (c.coalesce_max_requests == 0) c.coalesce_max_requests = 256;
    if (c.coalesce_max_requests < 2 || c.coalesce_max_requests > COALESCE_MAX_REQUESTS)
        throw Fail{HPCS_E_INVALID_ARG, "coalesce_max_requests must lie in [2, 1024]"};
    if (c.coalesce_max_batch_keys == 0) c.coalesce_max_batch_keys = int64_t(1) << 22;
    if (c.coalesce_max_batch_keys < 0 || c.coalesce_max_batch_keys >= (int64_t(1) << 31))
        throw Fail{HPCS_E_INVALID_ARG, "coalesce_max_batch_keys must lie in [0, 2^31)"};
    if (c.mode == HPCS_SERVE_ALONE) c.slots = 1;
    return c;
}

inline void validate(const hpcs_request *reqs, int64_t nreq)
{
    double prev = 0.0;
    for (int64_t i = 0; i < nreq; ++i) {
        const hpcs_request &r = reqs[i];
        if (r.op != HPCS_OP_SORT && r.op != HPCS_OP_TOPK) throw Fail{HPCS_E_INVALID_ARG, "unknown op"};
        if (r.dtype < HPCS_DTYPE_U32 || r.dtype > HPCS_DTYPE_F64) throw Fail{HPCS_E_INVALID_ARG, "unknown dtype"};
        if (r.n < 1 || r.n >= (int64_t(1) << 31)) throw Fail{HPCS_E_INVALID_ARG, "n must lie in [1, 2^31)"};
        if (r.op == HPCS_OP_TOPK && (r.k < 1 || r.k > r.n)) throw Fail{HPCS_E_INVALID_ARG, "k must lie in [1, n]"};
        if (r.arrival_s < prev) throw Fail{HPCS_E_INVALID_ARG, "arrivals must be non-decreasing"};
        prev = r.arrival_s;
    }
}

// Requests or batches enqueued on one slot, oldest first.
struct InFlight {
    std::vector<int64_t> idx;
    size_t off, bytes; // arena block
    int token;
    int64_t keys;
};

struct Slot {
    std::vector<int> spare; // free completion tokens
    std::deque<InFlight> inflight;
    std::mutex m;
    std::condition_variable cv;
    bool submit_done = false;
    int depth = 0; // guarded by Server::mu
};

template <class Exec>
class Server {
  public:
    Server(Exec &ex, const hpcs_serve_config &cfg, const hpcs_request *reqs, int64_t nreq,
           hpcs_request_result *res)
        : ex_(ex), cfg_(cfg), reqs_(reqs), res_(res), nreq_(nreq)
    {
    }

    // Run every request; on failure returns the first error's code and
    // message (and leaves device memory allocated if work may still run).
    int32_t run(hpcs_serve_stats *stats, std::string *err)
    {
        bool started = false;
        try {
            plan();
            ex_.setup(cfg_.slots, nmax_, batch_keys_cap(), batch_requests_cap(), arena_bytes(), nreq_);
            arena_.reset(ex_.arena_base(), arena_bytes());
            for (int s = 0; s < cfg_.slots; ++s) {
                slots_.emplace_back(new Slot());
                for (int t = 0; t < SLOT_DEPTH; ++t) slots_.back()->spare.push_back(t);
            }
            started = true;
            serve();
            if (!first_error_.empty()) throw Fail{first_code_, first_error_};
            std::vector<uint64_t> sums(size_t(std::max<int64_t>(nreq_, 1)));
            ex_.checksums(sums.data(), nreq_);
            double last = 0.0;
            int64_t keys = 0;
            for (int64_t i = 0; i < nreq_; ++i) {
                res_[i].checksum = sums[size_t(i)];
                last = std::max(last, res_[i].done_s);
                if (res_[i].status == HPCS_OK) keys += reqs_[i].n;
            }
            if (stats) {
                stats->wall_s = last;
                stats->keys = keys;
                stats->peak_bytes = peak_;
                stats->request_budget_bytes = budget_;
                stats->scratch_bytes_per_slot = scratch_;
                stats->slots = cfg_.slots;
                stats->batches = batches_;
                stats->coalesced_requests = coalesced_;
            }
        } catch (const Fail &e) {
            if (err) *err = e.msg;
            ex_.release(started && hung_);
            return e.code;
        }
        ex_.release(false);
        return HPCS_OK;
    }

  private:
    Exec &ex_;
    const hpcs_serve_config cfg_;
    const hpcs_request *reqs_;
    hpcs_request_result *res_;
    int64_t nreq_;

    int64_t budget_ = 0;  // bytes for request buffers (after slot scratch)
    int64_t scratch_ = 0; // bytes reserved per slot
    int64_t nmax_[2] = {0, 0};
    std::vector<char> rejected_; // requests that exceed the budget on their own
    hpcs::Arena arena_;          // guarded by mu_
    std::vector<std::unique_ptr<Slot>> slots_;

    std::mutex mu_; // admission state below
    std::condition_variable cv_;
    std::deque<int64_t> queue_;
    bool arrivals_done_ = false;
    int64_t reserved_ = 0, peak_ = 0, batches_ = 0, coalesced_ = 0;
    std::string first_error_;
    int32_t first_code_ = HPCS_OK;
    bool hung_ = false; // a slot's drain timed out
    Clock::time_point t0_;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }

    bool coalescable(const hpcs_request &r) const
    {
        return cfg_.coalesce_max_keys > 0 && r.op == HPCS_OP_SORT &&
               (r.dtype == HPCS_DTYPE_U32 || r.dtype == HPCS_DTYPE_F32) && r.n <= cfg_.coalesce_max_keys;
    }
    int64_t batch_keys_cap() const { return cfg_.coalesce_max_keys > 0 ? cfg_.coalesce_max_batch_keys : 0; }
    int batch_requests_cap() const { return cfg_.coalesce_max_keys > 0 ? cfg_.coalesce_max_requests : 0; }

    // Seconds a request (batch) may take once its slot's earlier ones are done.
    double timeout_for(int64_t keys) const
    {
        return cfg_.request_timeout_s > 0.0 ? cfg_.request_timeout_s : 5.0 + 20e-9 * double(keys);
    }

    // Slot scratch covers the largest accepted request of each key size and
    // the coalesced batches; a request whose buffers do not fit beside the
    // slots' scratch is rejected (HPCS_E_BUDGET), largest first, until the
    // rest fit.
    void plan()
    {
        rejected_.assign(size_t(nreq_), 0);
        std::vector<int64_t> order(static_cast<size_t>(nreq_));
        for (int64_t i = 0; i < nreq_; ++i) order[size_t(i)] = i;
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            return request_bytes(reqs_[a]) > request_bytes(reqs_[b]);
        });
        const int64_t seg = batch_keys_cap();
        for (size_t cut = 0;; ++cut) {
            nmax_[0] = nmax_[1] = 0;
            for (size_t j = cut; j < order.size(); ++j) {
                const hpcs_request &r = reqs_[order[j]];
                const int w = key_size(r.dtype) == 4 ? 0 : 1;
                nmax_[w] = std::max(nmax_[w], r.n);
            }
            const int64_t per_slot = std::max({ex_.slot_scratch(HPCS_DTYPE_U32, nmax_[0]),
                                               ex_.slot_scratch(HPCS_DTYPE_U64, nmax_[1]),
                                               ex_.slot_scratch(HPCS_DTYPE_U32, seg)});
            const int64_t left = cfg_.hbm_budget - per_slot * cfg_.slots;
            if (cut == order.size()) {
                if (left <= 0) throw Fail{HPCS_E_BUDGET, "the budget does not cover the slots' scratch"};
                scratch_ = per_slot;
                budget_ = left;
                break;
            }
            if (left > 0 && int64_t(request_bytes(reqs_[order[cut]])) <= left) {
                scratch_ = per_slot;
                budget_ = left;
                break;
            }
            rejected_[size_t(order[cut])] = 1;
        }
        // a coalesced batch must fit the request budget too
        if (cfg_.coalesce_max_keys > 0 &&
            int64_t(round_up(size_t(cfg_.coalesce_max_batch_keys) * 4) + size_t(cfg_.coalesce_max_requests) * 40 +
                    1024) > budget_)
            throw Fail{HPCS_E_BUDGET, "the budget does not cover one coalesced batch"};
    }

    // The arena never needs more than what the slots can hold in flight.
    size_t arena_bytes() const
    {
        size_t most = 0;
        for (int64_t i = 0; i < nreq_; ++i)
            if (!rejected_[size_t(i)]) most = std::max(most, request_bytes(reqs_[i]));
        if (cfg_.coalesce_max_keys > 0)
            most = std::max(most, round_up(size_t(cfg_.coalesce_max_batch_keys) * 4) +
                                      size_t(cfg_.coalesce_max_requests) * 40 + 1024);
        return std::min<size_t>(size_t(budget_), most * size_t(cfg_.slots) * SLOT_DEPTH);
    }

    void fail_all(int32_t code, const std::string &msg)
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (first_error_.empty()) {
                first_error_ = msg;
                first_code_ = code;
            }
        }
        cv_.notify_all();
        for (auto &sl : slots_) {
            std::lock_guard<std::mutex> lk(sl->m);
            sl->cv.notify_all();
        }
    }

    // Under mu_: the requests to admit next from the queue's front (one, or
    // the run of coalescable sorts there), with an arena block; false if the
    // block does not fit yet.
    bool take_front(std::vector<int64_t> &idx, size_t &off, size_t &bytes)
    {
        idx.assign(1, queue_.front());
        if (coalescable(reqs_[idx[0]])) {
            int64_t keys = reqs_[idx[0]].n;
            for (size_t q = 1; q < queue_.size() && int(idx.size()) < cfg_.coalesce_max_requests; ++q) {
                const int64_t j = queue_[q];
                if (rejected_[size_t(j)] || !coalescable(reqs_[j]) || keys + reqs_[j].n > cfg_.coalesce_max_batch_keys)
                    break;
                keys += reqs_[j].n;
                idx.push_back(j);
            }
        }
        bytes = batch_bytes(reqs_, idx.data(), int(idx.size()));
        off = arena_.alloc(bytes);
        if (off == SIZE_MAX && idx.size() > 1) { // the batch does not fit yet: its first request alone
            idx.resize(1);
            bytes = batch_bytes(reqs_, idx.data(), 1);
            off = arena_.alloc(bytes);
        }
        return off != SIZE_MAX;
    }

    void submitter(int s)
    {
        Slot &sl = *slots_[size_t(s)];
        std::vector<int32_t> status;
        try {
            for (;;) {
                std::vector<int64_t> idx;
                size_t bytes = 0, off = SIZE_MAX;
                {
                    std::unique_lock<std::mutex> lk(mu_);
                    cv_.wait(lk, [&] {
                        if (!first_error_.empty()) return true;
                        if (queue_.empty()) return arrivals_done_;
                        if (rejected_[size_t(queue_.front())]) return true;
                        if (sl.depth >= SLOT_DEPTH) return false;
                        return take_front(idx, off, bytes);
                    });
                    if (!first_error_.empty() || queue_.empty()) break;
                    if (rejected_[size_t(queue_.front())]) {
                        const int64_t i = queue_.front();
                        queue_.pop_front();
                        res_[i].admit_s = res_[i].enqueued_s = res_[i].done_s = now();
                        res_[i].slot = s;
                        res_[i].status = HPCS_E_BUDGET;
                        continue;
                    }
                    const double t = now();
                    for (int64_t i : idx) {
                        queue_.pop_front(); // idx is the queue's front, in order
                        res_[i].admit_s = t;
                        res_[i].slot = s;
                    }
                    reserved_ += int64_t(bytes);
                    peak_ = std::max(peak_, reserved_);
                    ++sl.depth;
                    ++batches_;
                    if (idx.size() > 1) coalesced_ += int64_t(idx.size());
                }
                int token;
                {
                    std::lock_guard<std::mutex> lk(sl.m);
                    token = sl.spare.back(); // depth <= SLOT_DEPTH keeps one free
                    sl.spare.pop_back();
                }
                status.assign(idx.size(), HPCS_OK);
                ex_.submit(s, token, reqs_, idx.data(), int(idx.size()), arena_.at(off), status.data());
                const double t = now();
                int64_t keys = 0;
                for (size_t k = 0; k < idx.size(); ++k) {
                    res_[idx[k]].status = status[k];
                    res_[idx[k]].enqueued_s = t;
                    keys += reqs_[idx[k]].n;
                }
                {
                    std::lock_guard<std::mutex> lk(sl.m);
                    sl.inflight.push_back({std::move(idx), off, bytes, token, keys});
                }
                sl.cv.notify_one();
            }
        } catch (const Fail &e) {
            fail_all(e.code, e.msg);
        }
        {
            std::lock_guard<std::mutex> lk(sl.m);
            sl.submit_done = true;
        }
        sl.cv.notify_one();
    }

    // Wait for f's completion token: polls, yielding for the first
    // millisecond and then sleeping 20 us between polls, until the deadline.
    void wait_done(int s, const InFlight &f)
    {
        const auto start = Clock::now();
        const double limit = timeout_for(f.keys);
        const auto deadline = start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(limit));
        for (;;) {
            if (ex_.poll(s, f.token)) return;
            const auto t = Clock::now();
            if (t > deadline) {
                const hpcs_request &r = reqs_[f.idx[0]];
                std::string what = "request " + std::to_string(f.idx[0]) + " (" +
                                   (r.op == HPCS_OP_SORT ? "sort" : "top-k") + " of " + std::to_string(r.n) +
                                   " keys";
                if (f.idx.size() > 1) what += ", first of " + std::to_string(f.idx.size()) + " coalesced";
                throw Fail{HPCS_E_HIP, what + ", slot " + std::to_string(s) + ") did not complete within " +
                                           std::to_string(limit) + " s"};
            }
            if (t - start < std::chrono::milliseconds(1)) std::this_thread::yield();
            else std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }

    void reaper(int s)
    {
        Slot &sl = *slots_[size_t(s)];
        double drain_s = 1.0; // after a failure: twice the last deadline, at least 1 s
        try {
            for (;;) {
                InFlight f;
                {
                    std::unique_lock<std::mutex> lk(sl.m);
                    sl.cv.wait(lk, [&] { return !sl.inflight.empty() || sl.submit_done; });
                    if (sl.inflight.empty()) return;
                    f = sl.inflight.front();
                }
                drain_s = std::max(1.0, 2.0 * timeout_for(f.keys));
                wait_done(s, f);
                const double t = now();
                for (int64_t i : f.idx) res_[i].done_s = t;
                {
                    std::lock_guard<std::mutex> lk(sl.m);
                    sl.inflight.pop_front();
                    sl.spare.push_back(f.token);
                }
                {
                    std::lock_guard<std::mutex> lk(mu_);
                    reserved_ -= int64_t(f.bytes);
                    arena_.release(f.off, f.bytes);
                    --sl.depth;
                }
                cv_.notify_all();
            }
        } catch (const Fail &e) {
            fail_all(e.code, e.msg);
            // drain: wait for what this slot enqueued so that nothing is freed
            // under it; if that does not finish either, keep the memory
            {
                std::unique_lock<std::mutex> lk(sl.m);
                sl.cv.wait(lk, [&] { return sl.submit_done; });
            }
            if (!ex_.drain(s, drain_s)) {
                std::lock_guard<std::mutex> lk(mu_);
                hung_ = true;
            }
        }
    }

    void serve()
    {
        t0_ = Clock::now();
        std::vector<std::thread> th;
        for (int s = 0; s < cfg_.slots; ++s) {
            th.emplace_back([this, s] { submitter(s); });
            th.emplace_back([this, s] { reaper(s); });
        }
        if (cfg_.mode == HPCS_SERVE_ALONE) {
            std::lock_guard<std::mutex> lk(mu_);
            for (int64_t i = 0; i < nreq_; ++i) queue_.push_back(i);
            arrivals_done_ = true;
            cv_.notify_all();
        } else {
            for (int64_t i = 0; i < nreq_; ++i) {
                const auto due = t0_ + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(reqs_[i].arrival_s));
                for (;;) {
                    const auto left = due - Clock::now();
                    if (left <= Clock::duration::zero()) break;
                    if (left > std::chrono::microseconds(300))
                        std::this_thread::sleep_for(left - std::chrono::microseconds(200));
                    else
                        std::this_thread::yield();
                }
                std::lock_guard<std::mutex> lk(mu_);
                if (!first_error_.empty()) break;
                queue_.push_back(i);
                cv_.notify_all();
            }
            std::lock_guard<std::mutex> lk(mu_);
            arrivals_done_ = true;
            cv_.notify_all();
        }
        for (auto &t : th) t.join();
    }
};

} // namespace hpcs_serve
