// rocprim_compare.hip -- libhpcsort's device-resident sort against rocPRIM's
// radix_sort_keys on the same keys, on the same GPU, in the same process.
// rocPRIM is only the yardstick here: the library neither links nor calls it.
//
// Keys: splitmix64 of the index (u32 keys: its low 32 bits), generated on the
// device.  Per (key type, n): rocPRIM sorts out of place (its input stays
// untouched); libhpcsort sorts a copy in place (the copy is timed on its own
// and subtracted).  Two warm-up calls, then the median of 10 timed calls
// between HIP events on one stream; the two outputs are compared word for
// word on the device.  libhpcsort uses the tuning table given as argv[1]
// (default hpcsharp_amd/tune_gfx950.txt, relative to the working directory).
//
// Build: python -m hpcsharp_amd.build --comparator  (writes hpcsharp_amd/rocprim_compare)
// Run:   ./hpcsharp_amd/rocprim_compare  (from the repository root)
#include <hip/hip_runtime.h>

#include <cstring> // rocPRIM's headers call memset without including it
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "hpcsort.h"

#define CHECK(x)                                                                                                       \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) {                                                                                        \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);                                 \
            return 1;                                                                                                  \
        }                                                                                                              \
    } while (0)
#define CHECK_LIB(x)                                                                                                   \
    do {                                                                                                               \
        int32_t s_ = (x);                                                                                              \
        if (s_ != 0) {                                                                                                 \
            std::printf("libhpcsort error %d (%s) at line %d\n", s_, hpcs_last_error(), __LINE__);                     \
            return 1;                                                                                                  \
        }                                                                                                              \
    } while (0)

template <typename K>
__global__ void k_gen(K *p, size_t n, uint64_t seed)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        uint64_t z = seed + (uint64_t(i) + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        p[i] = K(z ^ (z >> 31));
    }
}

template <typename K>
__global__ void k_diff(const K *a, const K *b, size_t n, unsigned long long *bad)
{
    unsigned long long c = 0;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        c += a[i] != b[i];
    if (c) atomicAdd(bad, c);
}

static float median(std::vector<float> v)
{
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

template <typename K>
static int run(hpcs_ctx *ctx, hipStream_t st, int lg)
{
    const size_t n = size_t(1) << lg;
    K *in = nullptr, *out = nullptr, *work = nullptr;
    unsigned long long *bad = nullptr;
    CHECK(hipMalloc(&in, n * sizeof(K)));
    CHECK(hipMalloc(&out, n * sizeof(K)));
    CHECK(hipMalloc(&work, n * sizeof(K)));
    CHECK(hipMalloc(&bad, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_gen<K>, dim3(4096), dim3(256), 0, st, in, n, 0x5EED0002ull + uint64_t(lg));
    CHECK(hipGetLastError());
    size_t temp_bytes = 0;
    CHECK(rocprim::radix_sort_keys(nullptr, temp_bytes, in, out, n, 0, int(8 * sizeof(K)), st));
    void *temp = nullptr;
    CHECK(hipMalloc(&temp, temp_bytes));
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    std::vector<float> t_prim, t_copy, t_lib;
    for (int r = 0; r < 12; ++r) {
        float ms = 0.f;
        CHECK(hipEventRecord(a, st));
        CHECK(rocprim::radix_sort_keys(temp, temp_bytes, in, out, n, 0, int(8 * sizeof(K)), st));
        CHECK(hipEventRecord(b, st));
        CHECK(hipEventSynchronize(b));
        CHECK(hipEventElapsedTime(&ms, a, b));
        if (r >= 2) t_prim.push_back(ms);
        CHECK(hipEventRecord(a, st));
        CHECK(hipMemcpyAsync(work, in, n * sizeof(K), hipMemcpyDeviceToDevice, st));
        CHECK(hipEventRecord(b, st));
        CHECK(hipEventSynchronize(b));
        CHECK(hipEventElapsedTime(&ms, a, b));
        if (r >= 2) t_copy.push_back(ms);
        CHECK(hipEventRecord(a, st));
        CHECK(hipMemcpyAsync(work, in, n * sizeof(K), hipMemcpyDeviceToDevice, st));
        if (sizeof(K) == 4)
            CHECK_LIB(hpcs_dev_sort_radix_u32(ctx, reinterpret_cast<uint32_t *>(work), int64_t(n), st));
        else
            CHECK_LIB(hpcs_dev_sort_radix_u64(ctx, reinterpret_cast<uint64_t *>(work), int64_t(n), st));
        CHECK(hipEventRecord(b, st));
        CHECK(hipEventSynchronize(b));
        CHECK(hipEventElapsedTime(&ms, a, b));
        if (r >= 2) t_lib.push_back(ms);
    }
    CHECK(hipMemsetAsync(bad, 0, sizeof(unsigned long long), st));
    hipLaunchKernelGGL(k_diff<K>, dim3(4096), dim3(256), 0, st, out, work, n, bad);
    CHECK(hipGetLastError());
    unsigned long long nbad = 0;
    CHECK(hipMemcpyAsync(&nbad, bad, sizeof(nbad), hipMemcpyDeviceToHost, st));
    CHECK(hipStreamSynchronize(st));
    const float prim = median(t_prim), lib = median(t_lib) - median(t_copy);
    std::printf("u%d 2^%d: rocPRIM radix_sort_keys %.3f ms (%.2f Gkeys/s), libhpcsort %.3f ms (%.2f Gkeys/s), "
                "ratio %.2f, outputs %s\n",
                int(8 * sizeof(K)), lg, prim, double(n) / prim / 1e6, lib, double(n) / lib / 1e6, prim / lib,
                nbad == 0 ? "identical" : "DIFFER");
    std::fflush(stdout);
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    CHECK(hipFree(temp));
    CHECK(hipFree(bad));
    CHECK(hipFree(work));
    CHECK(hipFree(out));
    CHECK(hipFree(in));
    return nbad == 0 ? 0 : 2;
}

int main(int argc, char **argv)
{
    const char *table = argc > 1 ? argv[1] : "hpcsharp_amd/tune_gfx950.txt";
    hpcs_ctx *ctx = nullptr;
    CHECK_LIB(hpcs_ctx_create(0, &ctx));
    CHECK_LIB(hpcs_ctx_load_tuning(ctx, table));
    hipStream_t st;
    CHECK(hipStreamCreate(&st));
    int rc = 0;
    for (int lg : {12, 14, 16, 20, 24, 28}) rc |= run<uint32_t>(ctx, st, lg);
    for (int lg : {12, 14, 16, 20, 24, 28}) rc |= run<uint64_t>(ctx, st, lg);
    CHECK(hipStreamDestroy(st));
    CHECK_LIB(hpcs_ctx_destroy(ctx));
    return rc;
}
