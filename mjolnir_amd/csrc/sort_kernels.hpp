// Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8 bits per pass.
//
// This is the cell-list rebuild's sort of cell keys into a contiguous particle order (the
// reference does std::sort over (particle, cell) pairs, PeriodicGridCellList.hpp:260-274;
// its OpenMP build does a k-way merge sort, omp/sort.hpp:17-90).  Stability makes the
// order inside each cell the input order, so the neighbour lists and every force sum built
// from them are deterministic.
//
// Per pass: k_sort_hist counts digits per 2048-element tile into hist[digit][tile];
// k_sort_scan turns the digit-major table into exclusive offsets; k_sort_scatter ranks each
// element among equal digits of its tile with wave64 ballots and writes it.  Every kernel
// returns at once when *enable == 0, so a captured step can always launch them.
#pragma once
#include "device_common.hpp"

namespace mjg
{
constexpr int kSortThreads = 256;
constexpr int kSortItems   = 8;
constexpr int kSortTile    = kSortThreads * kSortItems;

__global__ void __launch_bounds__(kSortThreads)
k_sort_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
            uint32_t* __restrict__ hist, const int* __restrict__ enable)
{
    if(enable != nullptr && *enable == 0) {return;}
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * kSortTile;
#pragma unroll
    for(int k = 0; k < kSortItems; ++k)
    {
        const uint32_t idx = base + k * kSortThreads + threadIdx.x;
        if(idx < n) {atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);}
    }
    __syncthreads();
    hist[threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m entries with one block of 1024 threads.
__global__ void __launch_bounds__(1024)
k_sort_scan(uint32_t* __restrict__ hist, uint32_t m, const int* __restrict__ enable)
{
    if(enable != nullptr && *enable == 0) {return;}
    __shared__ uint32_t s[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t chunk = (m + 1023) / 1024;
    const uint32_t b = t * chunk;
    const uint32_t e = (b + chunk < m) ? b + chunk : m;
    uint32_t sum = 0;
    for(uint32_t i = b; i < e; ++i) {sum += hist[i];}
    s[t] = sum;
    __syncthreads();
    for(uint32_t off = 1; off < 1024; off <<= 1)
    {
        const uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum; // exclusive prefix of this thread's chunk
    for(uint32_t i = b; i < e; ++i)
    {
        const uint32_t v = hist[i];
        hist[i] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(kSortThreads)
k_sort_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
               uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n, int shift,
               const uint32_t* __restrict__ hist, const int* __restrict__ enable)
{
    if(enable != nullptr && *enable == 0) {return;}
    __shared__ uint32_t base[256];
    __shared__ uint32_t wcnt[kSortThreads / kWave][256];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t % kWave, w = t / kWave;
    base[t] = hist[t * gridDim.x + blockIdx.x];
    const uint64_t lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (kWave - lane));
    for(int k = 0; k < kSortItems; ++k)
    {
        for(uint32_t q = t; q < (kSortThreads / kWave) * 256; q += kSortThreads) {(&wcnt[0][0])[q] = 0;}
        __syncthreads();
        const uint32_t idx = blockIdx.x * kSortTile + k * kSortThreads + t;
        const bool valid = idx < n;
        const uint32_t key = valid ? kin[idx] : 0u;
        const uint32_t val = valid ? vin[idx] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for(int bit = 0; bit < 8; ++bit)
        {
            const bool set = (d >> bit) & 1u;
            const uint64_t bb = __ballot(set);
            peers &= set ? bb : ~bb;
        }
        const uint32_t rank = __popcll(peers & lt_mask);
        if(valid && rank == 0) {wcnt[w][d] = __popcll(peers);}
        __syncthreads();
        if(valid)
        {
            uint32_t off = base[d];
            for(uint32_t ww = 0; ww < w; ++ww) {off += wcnt[ww][d];}
            kout[off + rank] = key;
            vout[off + rank] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for(uint32_t ww = 0; ww < kSortThreads / kWave; ++ww) {add += wcnt[ww][t];}
        base[t] += add;
        __syncthreads();
    }
}

// Host side: sorts (keys, vals) in place in `passes` passes of 8 bits; the result ends in
// keys/vals (pointers are swapped with the alternates as needed).
struct SortBuffers
{
    uint32_t* keys = nullptr;
    uint32_t* vals = nullptr;
    uint32_t* keys_alt = nullptr;
    uint32_t* vals_alt = nullptr;
    uint32_t* hist = nullptr;
    uint32_t n = 0;
    uint32_t tiles = 0;
};

inline void radix_sort(SortBuffers& b, int key_bits, const int* enable, hipStream_t s)
{
    if(b.n == 0) {return;}
    const int passes = (key_bits + 7) / 8;
    for(int p = 0; p < passes; ++p)
    {
        hipLaunchKernelGGL(k_sort_hist, dim3(b.tiles), dim3(kSortThreads), 0, s, b.keys, b.n, 8 * p, b.hist, enable);
        hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, b.hist, 256u * b.tiles, enable);
        hipLaunchKernelGGL(k_sort_scatter, dim3(b.tiles), dim3(kSortThreads), 0, s, b.keys, b.vals,
                           b.keys_alt, b.vals_alt, b.n, 8 * p, b.hist, enable);
        std::swap(b.keys, b.keys_alt);
        std::swap(b.vals, b.vals_alt);
    }
}

} // mjg
