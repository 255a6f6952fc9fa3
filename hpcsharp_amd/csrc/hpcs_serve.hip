// hpcs_serve.hip -- the serving loop of libhpcsort.so.
//
// A trace of sort and radix top-k requests is replayed against one GPU.  The
// payload of a request (n keys from a seeded generator) is produced on the
// device when the request is admitted, so every buffer stays resident in HBM.
// Admission is first come, first served, against an HBM budget: a request is
// admitted when its slot has room and its buffers can be carved from the
// arena, one device allocation made up front and managed on the host
// (first fit, neighbours merged on release), so the hot path makes no
// allocator calls.  Each slot owns a stream and a context (whose scratch is
// reserved up front for the largest accepted request), a submitter thread
// that admits and enqueues, and a reaper thread that retires completions in
// order and returns their arena blocks; a slot keeps up to SLOT_DEPTH
// requests enqueued, so the host prepares the next while the GPU runs the
// previous ones, and the slots' streams overlap on the device.  A request
// that cannot fit beside the slots' scratch on its own is rejected with
// HPCS_E_BUDGET.
//
// The reference has no serving loop; this is the serving pattern built on its
// sort and top-k surface (SortRadix, RadixSortLsd.cs; SelectTopK,
// SelectionTopKRadix.cs:142-171).
#include "../../include/hpcsort.h"
#include "hpcs_arena.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#define HPCS_API extern "C" __attribute__((visibility("default")))

namespace {

constexpr uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}

// Key i of a payload is output i of splitmix64(seed); floating payloads are
// the same bits with an all-ones exponent turned finite.
template <typename K>
__global__ void k_gen_payload(K *__restrict__ out, uint64_t n, uint64_t seed, int dtype)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t x = mix64(seed + (i + 1) * GOLDEN);
        if (dtype == HPCS_DTYPE_F32) {
            uint32_t b = uint32_t(x);
            if (((b >> 23) & 0xFFu) == 0xFFu) b &= 0xFF7FFFFFu;
            x = b;
        } else if (dtype == HPCS_DTYPE_F64) {
            if (((x >> 52) & 0x7FFull) == 0x7FFull) x &= ~(1ull << 52);
        }
        out[i] = K(x);
    }
}

// Position-weighted checksum of a result: sum over i of
// mix64(bits_i + (i + 1) * GOLDEN), modulo 2^64.  Each workgroup writes its
// partial sum; one workgroup adds the partials (no atomics on one address).
constexpr int CHECK_BLOCKS = 1024;

template <typename K>
__global__ __launch_bounds__(256) void k_checksum_partial(const K *__restrict__ a, uint64_t n,
                                                          uint64_t *__restrict__ partial)
{
    __shared__ uint64_t s_w[4];
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    uint64_t s = 0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        s += mix64(uint64_t(a[i]) + (i + 1) * GOLDEN);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(1024) void k_checksum_final(const uint64_t *__restrict__ partial, int count,
                                                         unsigned long long *__restrict__ out)
{
    __shared__ uint64_t s_w[16];
    uint64_t s = threadIdx.x < unsigned(count) ? partial[threadIdx.x] : 0ull;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < 16; ++w) t += s_w[w];
        *out = t;
    }
}

thread_local std::string g_serve_error;

struct HipFail {
    int32_t code;
    std::string msg;
};

void hchk(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw HipFail{HPCS_E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

size_t key_size(int32_t dtype) { return (dtype == HPCS_DTYPE_U32 || dtype == HPCS_DTYPE_F32) ? 4 : 8; }

size_t round_up(size_t x) { return (x + 255) / 256 * 256; }

// Device bytes a request holds while in flight: its payload and, for top-k,
// the output.
size_t request_bytes(const hpcs_request &r)
{
    const size_t ks = key_size(r.dtype);
    return round_up(size_t(r.n) * ks) + (r.op == HPCS_OP_TOPK ? round_up(size_t(r.k) * ks) : 0);
}

// Scratch a slot's context holds for sorts and top-k of up to n keys.
int64_t slot_scratch(int32_t dtype, int64_t n)
{
    if (n <= 0) return 0;
    hpcs_plan_info pi{};
    int64_t s = 0;
    if (hpcs_plan_query(nullptr, dtype, n, 0, &pi) == HPCS_OK) s = pi.scratch_bytes;
    return std::max<int64_t>(s, int64_t(2 * size_t(n) * key_size(dtype)) + (int64_t(1) << 20));
}

int grid_for(uint64_t n) { return int(std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 2048)); }

using Clock = std::chrono::steady_clock;

// Requests a slot may have enqueued on its stream at once: the host enqueues
// the next while the GPU runs the previous ones.
constexpr int SLOT_DEPTH = 4;

struct InFlight {
    int64_t i;
    size_t off, bytes; // arena block
    hipEvent_t ev;
};

// One worker slot: a stream, a context, checksum partials, and the requests
// enqueued on it in order.  Its submitter thread admits and enqueues; its
// reaper thread waits for completions in order and releases the memory.
struct Slot {
    hipStream_t st = nullptr;
    hpcs_ctx *ctx = nullptr;
    uint64_t *partial = nullptr;
    std::vector<hipEvent_t> events; // SLOT_DEPTH, free ones in `spare`
    std::vector<hipEvent_t> spare;
    std::deque<InFlight> inflight;
    std::mutex m;
    std::condition_variable cv;
    bool submit_done = false;
    int depth = 0; // guarded by Server::mu
};

struct Server {
    int device = 0;
    int64_t budget = 0;  // bytes for request buffers (after slot scratch)
    int64_t scratch = 0; // bytes reserved per slot
    char *arena_mem = nullptr;
    hpcs::Arena arena; // guarded by mu
    std::vector<std::unique_ptr<Slot>> slots;
    unsigned long long *d_sums = nullptr;

    const hpcs_request *reqs = nullptr;
    hpcs_request_result *res = nullptr;
    int64_t nreq = 0;

    std::mutex mu; // admission state below
    std::condition_variable cv;
    std::deque<int64_t> queue;
    bool arrivals_done = false;
    int64_t reserved = 0;
    int64_t peak = 0;
    Clock::time_point t0;
    std::string first_error;
    std::vector<char> rejected; // requests that exceed the budget on their own

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void fail_all(const std::string &msg)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            if (first_error.empty()) first_error = msg;
        }
        cv.notify_all();
        for (auto &sl : slots) {
            std::lock_guard<std::mutex> lk(sl->m);
            sl->cv.notify_all();
        }
    }

    bool failed()
    {
        std::lock_guard<std::mutex> lk(mu);
        return !first_error.empty();
    }

    // Enqueue request i on slot s (no waiting); returns the call's status.
    int32_t enqueue(Slot &sl, int64_t i, char *block, hipEvent_t ev)
    {
        const hpcs_request &r = reqs[i];
        hipStream_t st = sl.st;
        const size_t ks = key_size(r.dtype);
        void *in = block, *out = nullptr;
        int32_t rc = HPCS_OK;
        if (r.op == HPCS_OP_TOPK) out = block + round_up(size_t(r.n) * ks);
        const uint64_t n = uint64_t(r.n);
        if (ks == 4)
            hipLaunchKernelGGL(k_gen_payload<uint32_t>, dim3(grid_for(n)), dim3(256), 0, st,
                               static_cast<uint32_t *>(in), n, r.seed, r.dtype);
        else
            hipLaunchKernelGGL(k_gen_payload<uint64_t>, dim3(grid_for(n)), dim3(256), 0, st,
                               static_cast<uint64_t *>(in), n, r.seed, r.dtype);
        hchk(hipGetLastError(), "payload generator");
        hpcs_ctx *c = sl.ctx;
        if (r.op == HPCS_OP_SORT) {
            switch (r.dtype) {
            case HPCS_DTYPE_U32: rc = hpcs_dev_sort_radix_u32(c, static_cast<uint32_t *>(in), r.n, st); break;
            case HPCS_DTYPE_U64: rc = hpcs_dev_sort_radix_u64(c, static_cast<uint64_t *>(in), r.n, st); break;
            case HPCS_DTYPE_F32: rc = hpcs_dev_sort_radix_f32(c, static_cast<float *>(in), r.n, st); break;
            default: rc = hpcs_dev_sort_radix_f64(c, static_cast<double *>(in), r.n, st); break;
            }
        } else {
            switch (r.dtype) {
            case HPCS_DTYPE_U32:
                rc = hpcs_dev_topk_u32(c, static_cast<uint32_t *>(in), r.n, r.k, static_cast<uint32_t *>(out), st);
                break;
            case HPCS_DTYPE_U64:
                rc = hpcs_dev_topk_u64(c, static_cast<uint64_t *>(in), r.n, r.k, static_cast<uint64_t *>(out), st);
                break;
            case HPCS_DTYPE_F32:
                rc = hpcs_dev_topk_f32(c, static_cast<float *>(in), r.n, r.k, static_cast<float *>(out), st);
                break;
            default:
                rc = hpcs_dev_topk_f64(c, static_cast<double *>(in), r.n, r.k, static_cast<double *>(out), st);
                break;
            }
        }
        const void *result = r.op == HPCS_OP_SORT ? in : out;
        const uint64_t rn = rc == HPCS_OK ? uint64_t(r.op == HPCS_OP_SORT ? r.n : r.k) : 0;
        const int g = int(std::min<uint64_t>(std::max<uint64_t>((rn + 1023) / 1024, 1), CHECK_BLOCKS));
        if (ks == 4)
            hipLaunchKernelGGL(k_checksum_partial<uint32_t>, dim3(g), dim3(256), 0, st,
                               static_cast<const uint32_t *>(result), rn, sl.partial);
        else
            hipLaunchKernelGGL(k_checksum_partial<uint64_t>, dim3(g), dim3(256), 0, st,
                               static_cast<const uint64_t *>(result), rn, sl.partial);
        hipLaunchKernelGGL(k_checksum_final, dim3(1), dim3(1024), 0, st, sl.partial, g, d_sums + i);
        hchk(hipGetLastError(), "checksum");
        hchk(hipEventRecord(ev, st), "hipEventRecord");
        return rc;
    }

    void submitter(int s)
    {
        Slot &sl = *slots[size_t(s)];
        try {
            hchk(hipSetDevice(device), "hipSetDevice");
            for (;;) {
                int64_t i = -1;
                size_t bytes = 0, off = SIZE_MAX;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] {
                        if (!first_error.empty()) return true;
                        if (queue.empty()) return arrivals_done;
                        if (rejected[size_t(queue.front())]) return true;
                        if (sl.depth >= SLOT_DEPTH) return false;
                        off = arena.alloc(request_bytes(reqs[queue.front()]));
                        return off != SIZE_MAX;
                    });
                    if (!first_error.empty() || queue.empty()) break;
                    i = queue.front();
                    queue.pop_front();
                    res[i].admit_s = now();
                    res[i].slot = s;
                    if (rejected[size_t(i)]) {
                        res[i].enqueued_s = res[i].done_s = res[i].admit_s;
                        res[i].status = HPCS_E_BUDGET;
                        continue;
                    }
                    bytes = request_bytes(reqs[i]);
                    reserved += int64_t(bytes);
                    peak = std::max(peak, reserved);
                    ++sl.depth;
                }
                hipEvent_t ev;
                {
                    std::lock_guard<std::mutex> lk(sl.m);
                    ev = sl.spare.back(); // depth <= SLOT_DEPTH keeps one free
                    sl.spare.pop_back();
                }
                res[i].status = enqueue(sl, i, arena.at(off), ev);
                res[i].enqueued_s = now();
                {
                    std::lock_guard<std::mutex> lk(sl.m);
                    sl.inflight.push_back({i, off, bytes, ev});
                }
                sl.cv.notify_one();
            }
        } catch (const HipFail &e) {
            fail_all(e.msg);
        }
        {
            std::lock_guard<std::mutex> lk(sl.m);
            sl.submit_done = true;
        }
        sl.cv.notify_one();
    }

    void reaper(int s)
    {
        Slot &sl = *slots[size_t(s)];
        try {
            hchk(hipSetDevice(device), "hipSetDevice");
            for (;;) {
                InFlight f;
                {
                    std::unique_lock<std::mutex> lk(sl.m);
                    sl.cv.wait(lk, [&] { return !sl.inflight.empty() || sl.submit_done; });
                    if (sl.inflight.empty()) return;
                    f = sl.inflight.front();
                }
                hipError_t q;
                while ((q = hipEventQuery(f.ev)) == hipErrorNotReady) std::this_thread::yield();
                hchk(q, "request completion");
                res[f.i].done_s = now();
                {
                    std::lock_guard<std::mutex> lk(sl.m);
                    sl.inflight.pop_front();
                    sl.spare.push_back(f.ev);
                }
                {
                    std::lock_guard<std::mutex> lk(mu);
                    reserved -= int64_t(f.bytes);
                    arena.release(f.off, f.bytes);
                    --sl.depth;
                }
                cv.notify_all();
            }
        } catch (const HipFail &e) {
            fail_all(e.msg);
            // drain: wait for what this slot enqueued so that nothing is freed under it
            std::unique_lock<std::mutex> lk(sl.m);
            sl.cv.wait(lk, [&] { return sl.submit_done; });
            (void)hipStreamSynchronize(sl.st);
        }
    }
};

void validate(const hpcs_request *reqs, int64_t nreq)
{
    double prev = 0.0;
    for (int64_t i = 0; i < nreq; ++i) {
        const hpcs_request &r = reqs[i];
        if (r.op != HPCS_OP_SORT && r.op != HPCS_OP_TOPK) throw HipFail{HPCS_E_INVALID_ARG, "unknown op"};
        if (r.dtype < HPCS_DTYPE_U32 || r.dtype > HPCS_DTYPE_F64) throw HipFail{HPCS_E_INVALID_ARG, "unknown dtype"};
        if (r.n < 1 || r.n >= (int64_t(1) << 31)) throw HipFail{HPCS_E_INVALID_ARG, "n must lie in [1, 2^31)"};
        if (r.op == HPCS_OP_TOPK && (r.k < 1 || r.k > r.n)) throw HipFail{HPCS_E_INVALID_ARG, "k must lie in [1, n]"};
        if (r.arrival_s < prev) throw HipFail{HPCS_E_INVALID_ARG, "arrivals must be non-decreasing"};
        prev = r.arrival_s;
    }
}

void release(Server &S)
{
    for (auto &p : S.slots) {
        Slot &sl = *p;
        if (sl.st) (void)hipStreamSynchronize(sl.st);
        if (sl.ctx) hpcs_ctx_destroy(sl.ctx);
        if (sl.st) (void)hipStreamDestroy(sl.st);
        for (auto ev : sl.events) (void)hipEventDestroy(ev);
        if (sl.partial) (void)hipFree(sl.partial);
    }
    S.slots.clear();
    if (S.arena_mem) (void)hipFree(S.arena_mem);
    if (S.d_sums) (void)hipFree(S.d_sums);
    S.arena_mem = nullptr;
    S.d_sums = nullptr;
}

} // namespace

HPCS_API const char *hpcs_serve_last_error(void) { return g_serve_error.c_str(); }

HPCS_API int32_t hpcs_serve_run(int32_t device, int32_t slots, int64_t hbm_budget, const hpcs_request *reqs,
                                int64_t nreq, int32_t mode, hpcs_request_result *results, hpcs_serve_stats *stats)
{
    Server S;
    try {
        if (!reqs || !results || nreq < 0) throw HipFail{HPCS_E_INVALID_ARG, "requests or results are null"};
        if (slots < 1 || slots > 64) throw HipFail{HPCS_E_INVALID_ARG, "slots must lie in [1, 64]"};
        if (mode != HPCS_SERVE_REALTIME && mode != HPCS_SERVE_ALONE) throw HipFail{HPCS_E_INVALID_ARG, "unknown mode"};
        if (hbm_budget <= 0) throw HipFail{HPCS_E_INVALID_ARG, "budget must be positive"};
        validate(reqs, nreq);
        if (mode == HPCS_SERVE_ALONE) slots = 1;
        std::memset(results, 0, sizeof(hpcs_request_result) * size_t(nreq));
        if (stats) std::memset(stats, 0, sizeof(*stats));
        hchk(hipSetDevice(device), "hipSetDevice");
        S.device = device;
        S.reqs = reqs;
        S.res = results;
        S.nreq = nreq;
        // Slot scratch covers the largest accepted request of each key size; a
        // request whose payload does not fit beside the slots' scratch is
        // rejected (HPCS_E_BUDGET), largest first, until the rest fit.
        S.rejected.assign(size_t(nreq), 0);
        std::vector<int64_t> order(static_cast<size_t>(nreq));
        for (int64_t i = 0; i < nreq; ++i) order[size_t(i)] = i;
        std::sort(order.begin(), order.end(),
                  [&](int64_t a, int64_t b) { return request_bytes(reqs[a]) > request_bytes(reqs[b]); });
        int64_t nmax[2] = {0, 0};
        for (size_t cut = 0;; ++cut) {
            nmax[0] = nmax[1] = 0;
            for (size_t j = cut; j < order.size(); ++j) {
                const hpcs_request &r = reqs[order[j]];
                const int w = key_size(r.dtype) == 4 ? 0 : 1;
                nmax[w] = std::max(nmax[w], r.n);
            }
            const int64_t per_slot = std::max(slot_scratch(HPCS_DTYPE_U32, nmax[0]), slot_scratch(HPCS_DTYPE_U64, nmax[1]));
            const int64_t left = hbm_budget - per_slot * slots;
            if (cut == order.size()) {
                if (left <= 0) throw HipFail{HPCS_E_BUDGET, "the budget does not cover the slots' scratch"};
                S.scratch = per_slot;
                S.budget = left;
                break;
            }
            if (left > 0 && int64_t(request_bytes(reqs[order[cut]])) <= left) {
                S.scratch = per_slot;
                S.budget = left;
                break;
            }
            S.rejected[size_t(order[cut])] = 1;
        }
        for (int s = 0; s < slots; ++s) {
            S.slots.emplace_back(new Slot());
            Slot &sl = *S.slots.back();
            int32_t rc = hpcs_ctx_create(device, &sl.ctx);
            if (rc != HPCS_OK) throw HipFail{rc, std::string("context: ") + hpcs_last_error()};
            if (nmax[0] && (rc = hpcs_ctx_reserve(sl.ctx, HPCS_DTYPE_U32, nmax[0])) != HPCS_OK)
                throw HipFail{rc, std::string("reserve: ") + hpcs_last_error()};
            if (nmax[1] && (rc = hpcs_ctx_reserve(sl.ctx, HPCS_DTYPE_U64, nmax[1])) != HPCS_OK)
                throw HipFail{rc, std::string("reserve: ") + hpcs_last_error()};
            hchk(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking), "hipStreamCreate");
            hchk(hipMalloc(&sl.partial, sizeof(uint64_t) * CHECK_BLOCKS), "hipMalloc partials");
            for (int d = 0; d < SLOT_DEPTH; ++d) {
                hipEvent_t ev;
                hchk(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
                sl.events.push_back(ev);
                sl.spare.push_back(ev);
            }
        }
        // the arena never needs more than what the slots can hold in flight
        size_t most = 0;
        for (int64_t i = 0; i < nreq; ++i)
            if (!S.rejected[size_t(i)]) most = std::max(most, request_bytes(reqs[i]));
        const size_t arena_bytes = std::min<size_t>(size_t(S.budget), most * size_t(slots) * SLOT_DEPTH);
        if (arena_bytes) hchk(hipMalloc(&S.arena_mem, arena_bytes), "hipMalloc arena");
        S.arena.reset(S.arena_mem, arena_bytes);
        hchk(hipMalloc(&S.d_sums, sizeof(unsigned long long) * size_t(std::max<int64_t>(nreq, 1))), "hipMalloc sums");
        hchk(hipDeviceSynchronize(), "hipDeviceSynchronize");

        S.t0 = Clock::now();
        std::vector<std::thread> th;
        for (int s = 0; s < slots; ++s) {
            th.emplace_back([&S, s] { S.submitter(s); });
            th.emplace_back([&S, s] { S.reaper(s); });
        }
        if (mode == HPCS_SERVE_ALONE) {
            std::lock_guard<std::mutex> lk(S.mu);
            for (int64_t i = 0; i < nreq; ++i) S.queue.push_back(i);
            S.arrivals_done = true;
            S.cv.notify_all();
        } else {
            for (int64_t i = 0; i < nreq; ++i) {
                const auto due = S.t0 + std::chrono::duration_cast<Clock::duration>(
                                            std::chrono::duration<double>(reqs[i].arrival_s));
                for (;;) {
                    const auto left = due - Clock::now();
                    if (left <= Clock::duration::zero()) break;
                    if (left > std::chrono::microseconds(300)) std::this_thread::sleep_for(left - std::chrono::microseconds(200));
                    else std::this_thread::yield();
                }
                std::lock_guard<std::mutex> lk(S.mu);
                S.queue.push_back(i);
                S.cv.notify_all();
            }
            std::lock_guard<std::mutex> lk(S.mu);
            S.arrivals_done = true;
            S.cv.notify_all();
        }
        for (auto &t : th) t.join();
        if (!S.first_error.empty()) throw HipFail{HPCS_E_HIP, S.first_error};
        std::vector<unsigned long long> sums(size_t(std::max<int64_t>(nreq, 1)));
        hchk(hipMemcpy(sums.data(), S.d_sums, sizeof(unsigned long long) * sums.size(), hipMemcpyDeviceToHost),
             "checksums");
        double last = 0.0;
        int64_t keys = 0;
        for (int64_t i = 0; i < nreq; ++i) {
            results[i].checksum = sums[size_t(i)];
            last = std::max(last, results[i].done_s);
            if (results[i].status == HPCS_OK) keys += reqs[i].n;
        }
        if (stats) {
            stats->wall_s = last;
            stats->keys = keys;
            stats->peak_bytes = S.peak;
            stats->request_budget_bytes = S.budget;
            stats->scratch_bytes_per_slot = S.scratch;
            stats->slots = slots;
        }
    } catch (const HipFail &e) {
        g_serve_error = e.msg;
        release(S);
        return e.code;
    }
    release(S);
    return HPCS_OK;
}
