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
// launches, in microseconds and TB/s (2 GiB moved per launch).  Two
// read-only patterns give the ceiling of the counting kernels (1 GiB read per
// launch): one workgroup per 4096-key tile in 16-byte loads, as the count
// kernels read, and a grid-stride loop over the resident workgroups.
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

// Read-only: workgroup b reads tile b (4096 keys) in 16-byte vectors and
// writes one word (the XOR of its keys) to out[b].
__global__ __launch_bounds__(BLOCK) void k_read_tile(const uint4 *__restrict__ in, uint32_t *__restrict__ out)
{
    const size_t base = size_t(blockIdx.x) * BLOCK * 4;
    uint4 v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = in[base + i * BLOCK + threadIdx.x];
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) x ^= v[i].x ^ v[i].y ^ v[i].z ^ v[i].w;
    for (int off = 32; off > 0; off >>= 1) x ^= __shfl_xor(x, off);
    __shared__ uint32_t s[BLOCK / 64];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = s[0] ^ s[1] ^ s[2] ^ s[3];
}

// Read-only, grid-stride: thread g reads vectors g, g + stride, ... (four in
// flight per round) and writes the XOR of its keys to out[g].
__global__ __launch_bounds__(BLOCK) void k_read_stride(const uint4 *__restrict__ in, size_t nvec,
                                                       uint32_t *__restrict__ out)
{
    const size_t g = size_t(blockIdx.x) * BLOCK + threadIdx.x, stride = size_t(gridDim.x) * BLOCK;
    uint32_t x = 0;
    for (size_t v0 = g; v0 < nvec; v0 += 4 * stride) {
        uint4 v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) v[i] = v0 + i * stride < nvec ? in[v0 + i * stride] : make_uint4(0, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) x ^= v[i].x ^ v[i].y ^ v[i].z ^ v[i].w;
    }
    out[g] = x;
}

// Counting: the tile read of k_read_tile, each key's low byte counted with an
// LDS atomic, and the 256 counts written to out[b * 256 ...], the work of the
// sort's per-tile digit counts.  With FLAG, the source buffer is picked by a
// word read first (as the sort's kernels read their pass flags), so the key
// loads wait for that load.
template <bool FLAG>
__global__ __launch_bounds__(BLOCK) void k_count_tile(const uint4 *__restrict__ in, const uint4 *__restrict__ alt,
                                                      const uint32_t *__restrict__ flag, uint32_t *__restrict__ out)
{
    __shared__ uint32_t sh[256];
    const uint4 *src = in;
    if (FLAG) src = flag[0] ? alt : in;
    const size_t base = size_t(blockIdx.x) * BLOCK * 4;
    uint4 v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = src[base + i * BLOCK + threadIdx.x];
    sh[threadIdx.x] = 0u;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        atomicAdd(&sh[v[i].x & 255u], 1u);
        atomicAdd(&sh[v[i].y & 255u], 1u);
        atomicAdd(&sh[v[i].z & 255u], 1u);
        atomicAdd(&sh[v[i].w & 255u], 1u);
    }
    __syncthreads();
    out[size_t(blockIdx.x) * 256 + threadIdx.x] = sh[threadIdx.x];
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
    const char *names[] = {"copy", "runs of 8 keys", "runs of 16 keys", "runs of 32 keys", "runs of 64 keys",
                           "read, tile/wg", "read, stride", "count, tile/wg", "count, flag first"};
    int dev = 0, cus = 0;
    CHECK(hipGetDevice(&dev));
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // resident workgroups of the grid-stride read: 8 per CU (its out[] has
    // cus * 8 * BLOCK words, far below n)
    const uint32_t stride_grid = uint32_t(cus) * 8;
    const uint4 *in4 = reinterpret_cast<const uint4 *>(in);
    // the flag word (0: read `in`) sits at the end of `out`, past the count rows
    uint32_t *flag = out + (n - 64);
    CHECK(hipMemset(flag, 0, 64 * sizeof(uint32_t)));
    for (int k = 0; k < 9; ++k) {
        float best = 1e30f;
        for (int r = 0; r < 11; ++r) {
            CHECK(hipEventRecord(a));
            switch (k) {
            case 0: hipLaunchKernelGGL(k_copy, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in, out); break;
            case 1: hipLaunchKernelGGL(k_runs<8>, dim3(uint32_t(n / 2048)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 2048)); break;
            case 2: hipLaunchKernelGGL(k_runs<16>, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 4096)); break;
            case 3: hipLaunchKernelGGL(k_runs<32>, dim3(uint32_t(n / 8192)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 8192)); break;
            case 4: hipLaunchKernelGGL(k_runs<64>, dim3(uint32_t(n / 16384)), dim3(BLOCK), 0, 0, in, out, uint32_t(n / 16384)); break;
            case 5: hipLaunchKernelGGL(k_read_tile, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in4, out); break;
            case 6: hipLaunchKernelGGL(k_read_stride, dim3(stride_grid), dim3(BLOCK), 0, 0, in4, n / 4, out); break;
            case 7: hipLaunchKernelGGL(k_count_tile<false>, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in4, in4, flag, out); break;
            default: hipLaunchKernelGGL(k_count_tile<true>, dim3(uint32_t(n / 4096)), dim3(BLOCK), 0, 0, in4, in4, flag, out); break;
            }
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(b));
            CHECK(hipEventSynchronize(b));
            float ms = 0.f;
            CHECK(hipEventElapsedTime(&ms, a, b));
            if (r > 0 && ms < best) best = ms; // the first launch is a warm-up
        }
        const double bytes = double((k >= 5 ? 1 : 2) * n * sizeof(uint32_t));
        std::printf("%-16s %8.1f us  %5.2f TB/s\n", names[k], best * 1e3, bytes / (best * 1e-3) / 1e12);
    }
    CHECK(hipFree(in));
    CHECK(hipFree(out));
    return 0;
}
