// scatter_ceiling.hip -- what the memory system allows for the write
// pattern of a radix scatter pass, with no ranking work at all.
//
// 2^28 uint32 keys (1 GiB in, 1 GiB out).  Each workgroup of 256 threads
// reads one tile of 256 * R keys contiguously and writes it as 256 runs of R
// keys: run d of tile t lands at out[d * tiles * R + t * R ...], the
// bin-major, tile-minor layout of an LSD pass whose digits split every tile
// evenly.  Tiles are dealt XCD-contiguously, as the sort deals them.  R = 16
// is the mean run of an 8-bit pass over 4096-key tiles.  A plain copy of the
// same bytes is the reference.  Prints one line per pattern: best of 10
// launches, in microseconds and TB/s (2 GiB moved per launch).
//
// Build: python -m hpcsharp_amd.build (writes hpcsharp_amd/scatter_ceiling)
// Run:   ./hpcsharp_amd/scatter_ceiling
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                                                                       \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) {                                                                                        \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);                                 \
            return 1;                                                                                                  \
        }                                                                                                              \
    } while (0)

constexpr int BLOCK = 256;

// workgroup b -> tile: workgroups are dealt round-robin to the 8 XCDs, so
// XCD x = b % 8 gets the contiguous tile range x * g / 8 ...
__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t g)
{
    const uint32_t x = b % 8, i = b / 8, per = g / 8, rem = g % 8;
    return x * per + (x < rem ? x : rem) + i;
}

__global__ __launch_bounds__(BLOCK) void k_copy(const uint32_t *__restrict__ in, uint32_t *__restrict__ out)
{
    constexpr int ITEMS = 16;
    const size_t base = size_t(blockIdx.x) * BLOCK * ITEMS;
    uint32_t v[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) v[i] = in[base + i * BLOCK + threadIdx.x];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) out[base + i * BLOCK + threadIdx.x] = v[i];
}

template <int R>
__global__ __launch_bounds__(BLOCK) void k_runs(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                uint32_t tiles)
{
    constexpr int TILE = 256 * R, ITEMS = TILE / BLOCK;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const size_t base = size_t(t) * TILE;
    uint32_t v[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) v[i] = in[base + i * BLOCK + threadIdx.x];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t j = uint32_t(i * BLOCK) + threadIdx.x; // slot j: run j / R, offset j % R
        out[size_t(j / R) * tiles * R + size_t(t) * R + (j % R)] = v[i];
    }
}

__global__ void k_init(uint32_t *p, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = uint32_t(i * 2654435761u);
}

int main()
{
    const size_t n = size_t(1) << 28;
    uint32_t *in = nullptr, *out = nullptr;
    CHECK(hipMalloc(&in, n * sizeof(uint32_t)));
    CHECK(hipMalloc(&out, n * sizeof(uint32_t)));
    hipLaunchKernelGGL(k_init, dim3(4096), dim3(256), 0, 0, in, n);
    CHECK(hipDeviceSynchronize());
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    const char *names[] = {"copy", "runs of 8 keys", "runs of 16 keys", "runs of 32 keys", "runs of 64 keys"};
    for (int k = 0; k < 5; ++k) {
        float best = 1e30f;
        for (int r = 0; r < 11; ++r) {
            CHECK(hipEventRecord(a));
            switch (k) {
            case 0: hipLaunchKernelGGL(k_copy, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in, out); break;
            case 1: hipLaunchKernelGGL(k_runs<8>, dim3(uint32_t(n / 2048)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 2048)); break;
            case 2: hipLaunchKernelGGL(k_runs<16>, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 4096)); break;
            case 3: hipLaunchKernelGGL(k_runs<32>, dim3(uint32_t(n / 8192)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 8192)); break;
            default: hipLaunchKernelGGL(k_runs<64>, dim3(uint32_t(n / 16384)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 16384)); break;
            }
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b));
            CHECK(hipEventSynchronize(b));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, a, b));
            if (r > 0 && ms < best) best = ms; // the first launch is a warm-up
        }
        std::printf("%-16s %8.1f us  %5.2f TB/s\n", names[k], best * 1e3, double(2 * n * sizeof(uint32_t)) / (best * 1e-3) / 1e12);
    }
    CHECK(hipFree(in));
    CHECK(hipFree(out));
    return 0;
}
