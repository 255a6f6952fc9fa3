// hpcs_serve.hip -- the serving loop of libhpcsort.so: the HIP executor of
// hpcs_serve_core.hpp.
//
// A trace of sort and radix top-k requests is replayed against one GPU.  The
// payload of a request (n keys from a seeded generator) is produced on the
// device when the request is admitted, so every buffer stays resident in HBM.
// Admission (hpcs_serve_core.hpp) is first come, first served, against an
// HBM budget: a request is admitted when its slot has room and its buffers
// can be carved from the arena, one device allocation made up front and
// managed on the host, so the hot path makes no allocator calls.  Each slot
// owns a stream and a context (whose scratch is reserved up front for the
// largest accepted request and for the coalesced batches), a submitter
// thread that admits and enqueues, and a reaper thread that retires
// completions in order, each within a deadline, and returns their arena
// blocks; the slots' streams overlap on the device.  Small sorts of 4-byte
// keys queued together may be coalesced: one payload launch, one segmented
// sort and one checksum launch for the whole batch.  A request that cannot
// fit beside the slots' scratch on its own is rejected with HPCS_E_BUDGET.
//
// The reference has no serving loop; this is the serving pattern built on its
// sort and top-k surface (SortRadix, RadixSortLsd.cs; SelectTopK,
// SelectionTopKRadix.cs:142-171).
#include "../../include/hpcsort.h"
#include "hpcs_serve_core.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
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

// A coalesced batch (BatchLayout): workgroup row y generates segment y's keys
// (request ids[y]'s payload); float payloads are written as their radix keys
// (sign flipped, negative keys inverted), so the unsigned segmented sort
// gives their radix order.
__global__ __launch_bounds__(256) void k_gen_batch(uint32_t *__restrict__ keys, const int64_t *__restrict__ begin,
                                                  const uint64_t *__restrict__ seeds,
                                                  const uint32_t *__restrict__ flt)
{
    const int y = blockIdx.y;
    const uint64_t b = uint64_t(begin[y]), n = uint64_t(begin[y + 1]) - b, seed = seeds[y];
    const bool f = flt[y] != 0;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t v = uint32_t(mix64(seed + (i + 1) * GOLDEN));
        if (f) {
            if (((v >> 23) & 0xFFu) == 0xFFu) v &= 0xFF7FFFFFu;
            v ^= (v >> 31) ? 0xFFFFFFFFu : 0x80000000u;
        }
        keys[b + i] = v;
    }
}

// One workgroup per segment of a sorted batch: the segment's checksum (its
// float keys mapped back from radix keys to their bits) into sums[ids[y]].
__global__ __launch_bounds__(256) void k_checksum_batch(const uint32_t *__restrict__ keys,
                                                       const int64_t *__restrict__ begin,
                                                       const int64_t *__restrict__ ids,
                                                       const uint32_t *__restrict__ flt,
                                                       unsigned long long *__restrict__ sums)
{
    __shared__ uint64_t s_w[4];
    const int y = blockIdx.x;
    const uint64_t b = uint64_t(begin[y]), n = uint64_t(begin[y + 1]) - b;
    const bool f = flt[y] != 0;
    uint64_t s = 0;
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
        uint32_t v = keys[b + i];
        if (f) v ^= (v >> 31) ? 0x80000000u : 0xFFFFFFFFu;
        s += mix64(uint64_t(v) + (i + 1) * GOLDEN);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) sums[ids[y]] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
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

using hpcs_serve::Fail;
using hpcs_serve::key_size;
using hpcs_serve::round_up;
using hpcs_serve::SLOT_DEPTH;

void hchk(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw Fail{HPCS_E_OOM, std::string(what) + ": out of device memory"};
        throw Fail{HPCS_E_HIP, std::string(what) + ": " + hipGetErrorString(e)};
    }
}

int grid_for(uint64_t n) { return int(std::min<uint64_t>(std::max<uint64_t>((n + 255) / 256, 1), 2048)); }

// The executor on one HIP device: per slot a stream, a context, checksum
// partials, SLOT_DEPTH completion events and pinned staging for the batch
// tables; the arena and the per-request checksums.
struct HipExec {
    struct Slot {
        hipStream_t st = nullptr;
        hpcs_ctx *ctx = nullptr;
        uint64_t *partial = nullptr;
        hipEvent_t ev[SLOT_DEPTH] = {};
        char *stage = nullptr; // pinned: SLOT_DEPTH tables of stage_bytes
    };
    int device = 0;
    std::vector<Slot> slots;
    char *arena_mem = nullptr;
    unsigned long long *d_sums = nullptr;
    size_t stage_bytes = 0;

    // Scratch a slot's context holds for sorts and top-k of up to n keys.
    int64_t slot_scratch(int32_t dtype, int64_t n)
    {
        if (n <= 0) return 0;
        hpcs_plan_info pi{};
        int64_t s = 0;
        if (hpcs_plan_query(nullptr, dtype, n, 0, &pi) == HPCS_OK) s = pi.scratch_bytes;
        return std::max<int64_t>(s, int64_t(2 * size_t(n) * key_size(dtype)) + (int64_t(1) << 20));
    }

    void setup(int nslots, const int64_t nmax[2], int64_t batch_keys, int batch_requests, size_t arena_bytes,
               int64_t nreq)
    {
        hchk(hipSetDevice(device), "hipSetDevice");
        stage_bytes = batch_requests ? round_up(size_t(batch_requests + 1) * 8) + 2 * round_up(size_t(batch_requests) * 8) +
                                           round_up(size_t(batch_requests) * 4)
                                     : 0;
        slots.resize(size_t(nslots));
        for (auto &sl : slots) {
            int32_t rc = hpcs_ctx_create(device, &sl.ctx);
            if (rc != HPCS_OK) throw Fail{rc, std::string("context: ") + hpcs_last_error()};
            if (nmax[0] && (rc = hpcs_ctx_reserve(sl.ctx, HPCS_DTYPE_U32, nmax[0])) != HPCS_OK)
                throw Fail{rc, std::string("reserve: ") + hpcs_last_error()};
            if (nmax[1] && (rc = hpcs_ctx_reserve(sl.ctx, HPCS_DTYPE_U64, nmax[1])) != HPCS_OK)
                throw Fail{rc, std::string("reserve: ") + hpcs_last_error()};
            if (batch_keys && (rc = hpcs_ctx_reserve_segments(sl.ctx, batch_keys, batch_requests)) != HPCS_OK)
                throw Fail{rc, std::string("reserve: ") + hpcs_last_error()};
            hchk(hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking), "hipStreamCreate");
            hchk(hipMalloc(&sl.partial, sizeof(uint64_t) * CHECK_BLOCKS), "hipMalloc partials");
            for (auto &e : sl.ev) hchk(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
            if (stage_bytes) hchk(hipHostMalloc(&sl.stage, stage_bytes * SLOT_DEPTH, hipHostMallocDefault), "hipHostMalloc");
        }
        if (arena_bytes) hchk(hipMalloc(&arena_mem, arena_bytes), "hipMalloc arena");
        hchk(hipMalloc(&d_sums, sizeof(unsigned long long) * size_t(std::max<int64_t>(nreq, 1))), "hipMalloc sums");
        hchk(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }

    char *arena_base() { return arena_mem; }

    // One request: payload, the sort or top-k call, its checksum.
    int32_t submit_one(Slot &sl, const hpcs_request &r, int64_t i, char *block)
    {
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
        return rc;
    }

    // A coalesced batch: its tables through the token's pinned staging, one
    // payload launch, one segmented sort, one checksum launch.
    int32_t submit_batch(Slot &sl, int token, const hpcs_request *reqs, const int64_t *idx, int count, char *block)
    {
        hipStream_t st = sl.st;
        const hpcs_serve::BatchLayout L = hpcs_serve::batch_layout(reqs, idx, count);
        char *stage = sl.stage + size_t(token) * stage_bytes;
        const size_t tb = L.total - L.off_begin;            // the tables' bytes in the block
        if (tb > stage_bytes) throw Fail{HPCS_E_HIP, "batch tables exceed the staging area"};
        int64_t *begin = reinterpret_cast<int64_t *>(stage);
        uint64_t *seeds = reinterpret_cast<uint64_t *>(stage + (L.off_seed - L.off_begin));
        int64_t *ids = reinterpret_cast<int64_t *>(stage + (L.off_ids - L.off_begin));
        uint32_t *flt = reinterpret_cast<uint32_t *>(stage + (L.off_flt - L.off_begin));
        int64_t at = 0, most = 0;
        for (int k = 0; k < count; ++k) {
            const hpcs_request &r = reqs[idx[k]];
            begin[k] = at;
            seeds[k] = r.seed;
            ids[k] = idx[k];
            flt[k] = r.dtype == HPCS_DTYPE_F32 ? 1u : 0u;
            at += r.n;
            most = std::max(most, r.n);
        }
        begin[count] = at;
        char *dt = block + L.off_begin;
        hchk(hipMemcpyAsync(dt, stage, tb, hipMemcpyHostToDevice, st), "batch tables");
        uint32_t *keys = reinterpret_cast<uint32_t *>(block);
        const int64_t *d_begin = reinterpret_cast<const int64_t *>(dt);
        const uint64_t *d_seeds = reinterpret_cast<const uint64_t *>(block + L.off_seed);
        const int64_t *d_ids = reinterpret_cast<const int64_t *>(block + L.off_ids);
        const uint32_t *d_flt = reinterpret_cast<const uint32_t *>(block + L.off_flt);
        const unsigned gx = unsigned(std::min<int64_t>(std::max<int64_t>((most + 1023) / 1024, 1), 64));
        hipLaunchKernelGGL(k_gen_batch, dim3(gx, unsigned(count)), dim3(256), 0, st, keys, d_begin, d_seeds, d_flt);
        hchk(hipGetLastError(), "batch payload generator");
        const int32_t rc = hpcs_dev_sort_segments_u32(sl.ctx, keys, L.keys, d_begin, count, st);
        if (rc == HPCS_OK) {
            hipLaunchKernelGGL(k_checksum_batch, dim3(unsigned(count)), dim3(256), 0, st, keys, d_begin, d_ids, d_flt,
                               d_sums);
            hchk(hipGetLastError(), "batch checksum");
        }
        return rc;
    }

    void submit(int s, int token, const hpcs_request *reqs, const int64_t *idx, int count, char *block,
                int32_t *status)
    {
        Slot &sl = slots[size_t(s)];
        if (count == 1) {
            status[0] = submit_one(sl, reqs[idx[0]], idx[0], block);
        } else {
            const int32_t rc = submit_batch(sl, token, reqs, idx, count, block);
            for (int k = 0; k < count; ++k) status[k] = rc;
        }
        hchk(hipEventRecord(sl.ev[token], sl.st), "hipEventRecord");
    }

    bool poll(int s, int token)
    {
        const hipError_t q = hipEventQuery(slots[size_t(s)].ev[token]);
        if (q == hipErrorNotReady) return false;
        hchk(q, "request completion");
        return true;
    }

    bool drain(int s, double seconds)
    {
        const auto end = hpcs_serve::Clock::now() +
                         std::chrono::duration_cast<hpcs_serve::Clock::duration>(std::chrono::duration<double>(seconds));
        for (;;) {
            const hipError_t q = hipStreamQuery(slots[size_t(s)].st);
            if (q != hipErrorNotReady) return true; // complete, or failed: nothing runs any more
            if (hpcs_serve::Clock::now() > end) return false;
            std::this_thread::sleep_for(std::chrono::microseconds(100));
        }
    }

    void checksums(uint64_t *out, int64_t nreq)
    {
        hchk(hipMemcpy(out, d_sums, sizeof(unsigned long long) * size_t(std::max<int64_t>(nreq, 1)),
                       hipMemcpyDeviceToHost),
             "checksums");
    }

    void release(bool leak)
    {
        for (auto &sl : slots) {
            if (!leak && sl.st) (void)hipStreamSynchronize(sl.st);
            if (!leak && sl.ctx) hpcs_ctx_destroy(sl.ctx);
            if (!leak && sl.st) (void)hipStreamDestroy(sl.st);
            if (!leak)
                for (auto e : sl.ev)
                    if (e) (void)hipEventDestroy(e);
            if (!leak && sl.partial) (void)hipFree(sl.partial);
            if (!leak && sl.stage) (void)hipHostFree(sl.stage);
        }
        slots.clear();
        if (!leak && arena_mem) (void)hipFree(arena_mem);
        if (!leak && d_sums) (void)hipFree(d_sums);
        arena_mem = nullptr;
        d_sums = nullptr;
    }
};

} // namespace

HPCS_API const char *hpcs_serve_last_error(void) { return g_serve_error.c_str(); }

HPCS_API int32_t hpcs_serve_run_config(int32_t device, const hpcs_serve_config *cfg, const hpcs_request *reqs,
                                       int64_t nreq, hpcs_request_result *results, hpcs_serve_stats *stats)
{
    HipExec ex;
    ex.device = device;
    try {
        if (!cfg) throw Fail{HPCS_E_INVALID_ARG, "config is null"};
        if (!reqs || !results || nreq < 0) throw Fail{HPCS_E_INVALID_ARG, "requests or results are null"};
        const hpcs_serve_config c = hpcs_serve::resolve_config(*cfg);
        hpcs_serve::validate(reqs, nreq);
        std::memset(results, 0, sizeof(hpcs_request_result) * size_t(nreq));
        if (stats) std::memset(stats, 0, sizeof(*stats));
        hpcs_serve::Server<HipExec> S(ex, c, reqs, nreq, results);
        std::string err;
        const int32_t rc = S.run(stats, &err);
        if (rc != HPCS_OK) g_serve_error = err;
        return rc;
    } catch (const Fail &e) {
        g_serve_error = e.msg;
        return e.code;
    }
}

HPCS_API int32_t hpcs_serve_run(int32_t device, int32_t slots, int64_t hbm_budget, const hpcs_request *reqs,
                                int64_t nreq, int32_t mode, hpcs_request_result *results, hpcs_serve_stats *stats)
{
    hpcs_serve_config c{};
    c.slots = slots;
    c.mode = mode;
    c.hbm_budget = hbm_budget;
    return hpcs_serve_run_config(device, &c, reqs, nreq, results, stats);
}
