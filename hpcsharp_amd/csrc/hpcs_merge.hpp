// hpcs_merge.hpp -- the memory-tight form of a keys-only sort: each half of
// the array is sorted in place by the LSD passes (scratch for half the keys),
// then the two sorted halves are merged back into the caller's buffer with
// half the keys of scratch.
//
// Halves L = keys[0, h) and R = keys[h, n), h = ceil(n / 2), nb = n - h <= h:
//   1. L is copied to scratch S[0, h).
//   2. k_merge_split: i = the number of L's keys among the first h outputs
//      (merge-path split, L's key first among equal keys), j = h - i.
//   3. k_merge phase 0 writes outputs [0, h) from S[0, h) and R: the outputs
//      lie in L's old place, which no source uses any more.
//   4. k_merge_tail copies R's unmerged keys R[j, nb) to S[0, nb - j), the
//      part of S whose keys step 3 consumed (nb - j <= h - j = i).
//   5. k_merge phase 1 writes outputs [h, n) from S[i, h) and S[0, nb - j):
//      both sources are in S, the outputs in R's old place.
// Keys compare by their radix key (the float / signed transforms), so the
// result is the radix sort's order; for keys without values the output is
// the same sorted array the LSD passes give.
#pragma once
#include "hpcs_kernels.hpp"

namespace hpcs {

struct MergeSplit {
    uint64_t i, j; // L's and R's keys among the first h outputs
};

// Number of a's keys among the first d outputs of the merge of a[0, na) and
// b[0, nb) (a's keys first among equal keys); the smallest i with
// a[i] > b[d - i - 1].
template <typename K, int KIND>
__device__ inline uint64_t merge_path_split(const K *a, uint64_t na, const K *b, uint64_t nb, uint64_t d)
{
    uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (key_enc<K, KIND>(a[mid]) <= key_enc<K, KIND>(b[d - mid - 1])) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <typename K, int KIND>
__global__ void k_merge_split(const K *__restrict__ l, uint64_t h, const K *__restrict__ r, uint64_t nb,
                              MergeSplit *__restrict__ sp)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint64_t i = merge_path_split<K, KIND>(l, h, r, nb, h);
        sp->i = i;
        sp->j = h - i;
    }
}

constexpr int MERGE_BLOCK = 256;
constexpr int MERGE_ITEMS = 8;
constexpr int MERGE_TILE = MERGE_BLOCK * MERGE_ITEMS; // outputs per workgroup

// Phase 0: out[0, h) = first h outputs of merge(s[0, h), r[0, nb)).
// Phase 1: out[0, nb) = merge(s[i, h), s[0, nb - j)) (the remaining keys).
// Workgroup w produces outputs [w * MERGE_TILE, ...): it finds its two
// splits, stages its slice of both runs in LDS and each thread merges
// MERGE_ITEMS outputs from its own split.  Spare workgroups return.
template <typename K, int KIND>
__global__ __launch_bounds__(MERGE_BLOCK) void k_merge(const K *__restrict__ s, uint64_t h, const K *__restrict__ r,
                                                       uint64_t nb, const MergeSplit *__restrict__ sp, int phase,
                                                       K *__restrict__ out)
{
    __shared__ K t[MERGE_TILE];
    __shared__ uint64_t s_cut[2];
    const K *a, *b;
    uint64_t na, nbb;
    if (phase == 0) {
        a = s, na = h, b = r, nbb = nb;
    } else {
        a = s + sp->i, na = h - sp->i, b = s, nbb = nb - sp->j;
    }
    const uint64_t total = phase == 0 ? h : na + nbb;
    const uint64_t ds = uint64_t(blockIdx.x) * MERGE_TILE;
    if (ds >= total) return; // workgroup-uniform
    const uint64_t de = ds + MERGE_TILE < total ? ds + MERGE_TILE : total;
    if (threadIdx.x < 2) s_cut[threadIdx.x] = merge_path_split<K, KIND>(a, na, b, nbb, threadIdx.x ? de : ds);
    __syncthreads();
    const uint64_t ia = s_cut[0], ib = s_cut[1];
    const int la = int(ib - ia), lb = int((de - ib) - (ds - ia)); // la + lb = de - ds <= MERGE_TILE
    const uint64_t ja = ds - ia;
    for (int q = threadIdx.x; q < la + lb; q += MERGE_BLOCK) t[q] = q < la ? a[ia + q] : b[ja + uint64_t(q - la)];
    __syncthreads();
    const int len = la + lb;
    const int dt = min(int(threadIdx.x) * MERGE_ITEMS, len);
    int lo = dt > lb ? dt - lb : 0, hi = dt < la ? dt : la;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (key_enc<K, KIND>(t[mid]) <= key_enc<K, KIND>(t[la + dt - mid - 1])) lo = mid + 1;
        else hi = mid;
    }
    int i = lo, j = dt - lo;
    for (int k = 0; k < MERGE_ITEMS && dt + k < len; ++k) {
        const bool take_a = j >= lb || (i < la && key_enc<K, KIND>(t[i]) <= key_enc<K, KIND>(t[la + j]));
        out[ds + uint64_t(dt + k)] = take_a ? t[i++] : t[la + j++];
    }
}

// One pass of the block-sort + merge sort (n up to a few times 10^5 keys,
// where launches, not bytes, set the time): src holds sorted runs of `run`
// keys (the last one may be shorter), and the runs of each pair (2p, 2p + 1)
// are merged into dst, stably (the left run's key first among equal keys),
// values (HAS_VALS) moving with their keys.  Workgroup w produces outputs
// [w * MERGE_TILE, ...), which lie in one pair since 2 * run is a multiple of
// MERGE_TILE; a last run without a partner is copied.  Keys compare by their
// radix key (KIND), so the order is the radix sort's.
template <typename K, bool HAS_VALS, int KIND>
__global__ __launch_bounds__(MERGE_BLOCK) void k_merge_runs(const K *__restrict__ src,
                                                            const uint32_t *__restrict__ vsrc, K *__restrict__ dst,
                                                            uint32_t *__restrict__ vdst, uint64_t n, uint64_t run)
{
    __shared__ K t[MERGE_TILE];
    __shared__ uint32_t tv[HAS_VALS ? MERGE_TILE : 1];
    __shared__ uint64_t s_cut[2];
    const uint64_t ds = uint64_t(blockIdx.x) * MERGE_TILE;
    if (ds >= n) return; // workgroup-uniform
    const uint64_t ps = ds / (2 * run) * (2 * run); // the pair's first key
    const uint64_t na = min(run, n - ps);
    const uint64_t nb = ps + na < n ? min(run, n - ps - na) : 0;
    const K *a = src + ps, *b = src + ps + na;
    const uint64_t d0 = ds - ps, d1 = min(d0 + uint64_t(MERGE_TILE), na + nb);
    if (threadIdx.x < 2) s_cut[threadIdx.x] = merge_path_split<K, KIND>(a, na, b, nb, threadIdx.x ? d1 : d0);
    __syncthreads();
    const uint64_t ia = s_cut[0], ib = s_cut[1];
    const int la = int(ib - ia), lb = int((d1 - ib) - (d0 - ia)); // la + lb = d1 - d0 <= MERGE_TILE
    const uint64_t ja = d0 - ia;
    for (int q = threadIdx.x; q < la + lb; q += MERGE_BLOCK) {
        const uint64_t at = q < la ? ps + ia + uint64_t(q) : ps + na + ja + uint64_t(q - la);
        t[q] = src[at];
        if constexpr (HAS_VALS) tv[q] = vsrc[at];
    }
    __syncthreads();
    const int len = la + lb;
    const int dt = min(int(threadIdx.x) * MERGE_ITEMS, len);
    int lo = dt > lb ? dt - lb : 0, hi = dt < la ? dt : la;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (key_enc<K, KIND>(t[mid]) <= key_enc<K, KIND>(t[la + dt - mid - 1])) lo = mid + 1;
        else hi = mid;
    }
    int i = lo, j = dt - lo;
    for (int k = 0; k < MERGE_ITEMS && dt + k < len; ++k) {
        const bool take_a = j >= lb || (i < la && key_enc<K, KIND>(t[i]) <= key_enc<K, KIND>(t[la + j]));
        const int q = take_a ? i++ : la + j++;
        dst[ds + uint64_t(dt + k)] = t[q];
        if constexpr (HAS_VALS) vdst[ds + uint64_t(dt + k)] = tv[q];
    }
}

// dst[0, n) = src[0, n) (different buffers).
template <typename K>
__global__ void k_copy_keys(K *__restrict__ dst, const K *__restrict__ src, uint64_t n)
{
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = src[q];
}

// R's unmerged keys r[j, nb) to s[0, nb - j).
template <typename K>
__global__ void k_merge_tail(const K *__restrict__ r, uint64_t nb, const MergeSplit *__restrict__ sp,
                             K *__restrict__ s)
{
    const uint64_t j = sp->j, cnt = nb - j;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    for (uint64_t q = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < cnt; q += stride) s[q] = r[j + q];
}

} // namespace hpcs
