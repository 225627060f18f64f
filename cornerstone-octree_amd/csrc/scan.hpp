// Device-wide exclusive scan of integers (reduce, scan partials, re-scan with offsets), input type In summed in
// type V. Used for the rebalance op scan (tree/csarray.hpp:380), the leaf layouts and the 64-bit count scan of the
// SFC assignment.
#pragma once

#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{
namespace scan
{

constexpr int kThreads = 256;
constexpr int kItems   = 16;
constexpr int kTile    = kThreads * kItems;
constexpr int kMaxPartialsPerThread = 16; // the single-block middle scan covers 1024*16 tiles
constexpr int kMidThreads           = 1024;

inline size_t numTiles(size_t n) { return (n + kTile - 1) / kTile; }

// bytes of workspace for a scan of n elements summed in V (partials + total)
template<class V = uint32_t>
inline size_t workspaceBytes(size_t n)
{
    return ((numTiles(n) + 1) * sizeof(V) + 255) & ~size_t(255);
}

// inclusive scan across one wave (64 lanes) with shuffles
template<class V>
__device__ inline V waveInclusiveScan(V v)
{
    const int lane = laneId();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1)
    {
        V t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

// exclusive scan across the block; returns this thread's exclusive prefix, total in *total
template<int Threads, class V>
__device__ inline V blockExclusiveScan(V v, V* total)
{
    constexpr int waves = Threads / kWave;
    __shared__ V shWave[waves];
    const int lane = laneId();
    const int wave = threadIdx.x / kWave;
    V incl         = waveInclusiveScan(v);
    if (lane == kWave - 1) shWave[wave] = incl;
    __syncthreads();
    V waveOff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < waves; ++w)
    {
        V s = shWave[w];
        if (w < wave) waveOff += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return waveOff + incl - v;
}

template<class In, class V>
__global__ __launch_bounds__(kThreads) void tileReduceKernel(const In* __restrict__ in, size_t n,
                                                             V* __restrict__ partials)
{
    size_t base = size_t(blockIdx.x) * kTile;
    V sum       = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i = base + size_t(k) * kThreads + threadIdx.x;
        if (i < n) sum += V(in[i]);
    }
    V total;
    blockExclusiveScan<kThreads>(sum, &total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// exclusive scan of the tile partials in place; partials[numTiles] = grand total
template<class V>
__global__ __launch_bounds__(kMidThreads) void partialsScanKernel(V* partials, int numTiles)
{
    V vals[kMaxPartialsPerThread];
    V local  = 0;
    int base = threadIdx.x * kMaxPartialsPerThread;
#pragma unroll
    for (int k = 0; k < kMaxPartialsPerThread; ++k)
    {
        int i   = base + k;
        vals[k] = i < numTiles ? partials[i] : V(0);
        local += vals[k];
    }
    V total;
    V run = blockExclusiveScan<kMidThreads>(local, &total);
#pragma unroll
    for (int k = 0; k < kMaxPartialsPerThread; ++k)
    {
        int i = base + k;
        if (i < numTiles) partials[i] = run;
        run += vals[k];
    }
    if (threadIdx.x == 0) partials[numTiles] = total;
}

// out[i] = init + partials[tile] + exclusive prefix within the tile; out may alias in.
// If writeTotalAt is non-null, *writeTotalAt = init + grand total.
template<class In, class V>
__global__ __launch_bounds__(kThreads) void tileScanKernel(const In* in, V* out, size_t n, const V* partials, V init,
                                                           V* writeTotalAt, int numTiles)
{
    size_t base = size_t(blockIdx.x) * kTile + size_t(threadIdx.x) * kItems; // blocked for the local scan
    V vals[kItems];
    V sum = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i = base + k;
        vals[k]  = i < n ? V(in[i]) : V(0);
        sum += vals[k];
    }
    V total;
    V run = blockExclusiveScan<kThreads>(sum, &total) + V(partials[blockIdx.x]) + init;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i = base + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
    if (writeTotalAt && blockIdx.x == 0 && threadIdx.x == 0) *writeTotalAt = init + V(partials[numTiles]);
}

template<class V>
__global__ void storeValueKernel(V* dst, V v)
{
    *dst = v;
}

// Exclusive scan of n values of in (type In) into out (type V; out may equal in when the types match). If totalOut
// is non-null it receives init + sum(in), e.g. out[n] for a scan of n+1 slots. Up to 1024*16 tiles (67 M elements).
template<class V, class In = V>
hipcstone_status exclusiveScan(hipStream_t s, const In* in, V* out, size_t n, V init, V* totalOut, void* ws,
                               size_t wsBytes)
{
    if (n == 0)
    {
        if (totalOut)
        {
            storeValueKernel<V><<<1, 1, 0, s>>>(totalOut, init);
            HC_CHECK_LAUNCH("storeValueKernel");
        }
        return HIPCSTONE_OK;
    }
    size_t tiles = numTiles(n);
    if (tiles > size_t(kMidThreads) * kMaxPartialsPerThread) return invalid("scan: too many elements");
    if (!ws || wsBytes < workspaceBytes<V>(n)) return invalid("scan: workspace too small");
    V* partials = static_cast<V*>(ws);
    tileReduceKernel<In, V><<<unsigned(tiles), kThreads, 0, s>>>(in, n, partials);
    HC_CHECK_LAUNCH("tileReduceKernel");
    partialsScanKernel<V><<<1, kMidThreads, 0, s>>>(partials, int(tiles));
    HC_CHECK_LAUNCH("partialsScanKernel");
    tileScanKernel<In, V><<<unsigned(tiles), kThreads, 0, s>>>(in, out, n, partials, init, totalOut, int(tiles));
    HC_CHECK_LAUNCH("tileScanKernel");
    return HIPCSTONE_OK;
}

} // namespace scan
} // namespace hipcstone
