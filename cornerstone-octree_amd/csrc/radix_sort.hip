// a4/a5: stable LSD radix sort (key + uint32 value) with a fused permutation gather.
//
// Replaces the reference's CUB DeviceRadixSort::SortPairs call (primitives/primitives_gpu.cu:322-356) and the
// gather kernels (:74-90). Result contract: ascending keys, ties in input order (what std::stable_sort gives on
// the CPU path, primitives/gather.hpp:42-67).
//
// Design (onesweep): one histogram kernel computes the digit histograms of all passes in a single read of the
// keys; then each 8-bit digit pass is a single kernel that
//   1. takes a tile id from an atomic counter (ids follow launch order, so every tile waits only on tiles that
//      have already started: the look-back always terminates),
//   2. ranks its 5888 keys stably with wave64 match-by-ballot (8 ballots per key) and per-wave LDS digit counters,
//   3. publishes its per-digit counts and resolves its global digit offsets by decoupled look-back
//      (one thread per digit, status word = 2-bit flag | 30-bit count),
//   4. reorders the tile in LDS (keys, then values in the same storage, so a block holds 8 bytes per key and three
//      blocks fit in a CU's LDS) and writes each digit run contiguously.
// The last pass optionally writes the permutation and gathers x,y,z,h through it. Gathering 4 separate arrays at
// random positions fetches 4 memory lines per particle; when the input is far from key order (more than n/8
// descents, counted by the histogram kernel), x,y,z,h are first packed into one 4-value record per particle so that
// the gather fetches one line per particle. Input already close to key order (a previous step's output) gathers the
// separate arrays directly, which is coalesced. Both read the same values, so the result does not depend on the
// choice.
#include <type_traits>

#include <mutex>
#include <string>

#include "common.hpp"
#include "runtime.hpp"
#include "sort.hpp"

namespace hipcstone
{

namespace sortcfg
{
constexpr int kThreads   = 256;
constexpr int kItems     = 23;
constexpr int kTile      = kThreads * kItems; // 5888 keys per tile
constexpr int kRadixBits = 8;
constexpr int kRadix     = 1 << kRadixBits;
constexpr int kWaves     = kThreads / kWave;
constexpr int kMaxPasses = 8;
constexpr int kHistBlocks = 1024;

constexpr uint32_t kFlagAggregate = 1u << 30;
constexpr uint32_t kFlagPrefix    = 2u << 30;
constexpr uint32_t kValueMask     = (1u << 30) - 1;
} // namespace sortcfg

using namespace sortcfg;

// Workspace layout (bytes, 256-aligned pieces):
//   hist    [kMaxPasses][kRadix] u32   digit histograms, then exclusive digit offsets
//   counter [kMaxPasses + 1] u32 (padded) tile-id counters, then the input's descent count
//   status  [kMaxPasses][numTiles][kRadix] u32 look-back words
//   packed  [n][4] x 8 bytes           x,y,z,h records (sort-gather workspaces only)
struct SortWorkspace
{
    uint32_t* hist;
    uint32_t* counter;
    uint32_t* status;
    void* packed;
    size_t bytes;
};

static size_t alignUp(size_t v) { return (v + 255) & ~size_t(255); }

static size_t numTiles(size_t n) { return (n + kTile - 1) / kTile; }

static SortWorkspace carveWorkspace(void* base, size_t n, int passes, bool packed)
{
    SortWorkspace w{};
    size_t off = 0;
    char* b    = static_cast<char*>(base);
    w.hist     = reinterpret_cast<uint32_t*>(b + off);
    off += alignUp(sizeof(uint32_t) * kMaxPasses * kRadix);
    w.counter = reinterpret_cast<uint32_t*>(b + off);
    off += alignUp(sizeof(uint32_t) * (kMaxPasses + 1));
    w.status = reinterpret_cast<uint32_t*>(b + off);
    off += alignUp(sizeof(uint32_t) * size_t(passes) * numTiles(n) * kRadix);
    if (packed)
    {
        w.packed = b + off;
        off += alignUp(n * 4 * sizeof(double));
    }
    w.bytes = off;
    return w;
}

static size_t workspaceBytes(size_t n, int passes, bool packed)
{
    return carveWorkspace(nullptr, n, passes, packed).bytes;
}

__device__ inline uint32_t atomicLoadRelaxed(const uint32_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void atomicStoreRelaxed(uint32_t* p, uint32_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Digit histograms of all passes [0, numPasses) in one read of the keys. hist must be zeroed.
// With descents set, also counts the positions i with keys[i] > keys[i + 1].
template<class K>
__global__ __launch_bounds__(kThreads) void histogramKernel(const K* __restrict__ keys, size_t n, int numPasses,
                                                            int beginBit, int endBit, uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ descents)
{
    __shared__ uint32_t sh[kMaxPasses * kRadix];
    __shared__ uint32_t shDescents;
    for (int i = threadIdx.x; i < kMaxPasses * kRadix; i += kThreads)
        sh[i] = 0;
    if (threadIdx.x == 0) shDescents = 0;
    __syncthreads();

    const int lane = laneId();
    size_t stride  = size_t(gridDim.x) * kThreads;
    for (size_t base = size_t(blockIdx.x) * kThreads; base < n; base += stride)
    {
        size_t i     = base + threadIdx.x;
        bool valid   = i < n;
        K key        = valid ? keys[i] : K(0);
        uint64_t act = ballot(valid);
        if (act == 0) continue;
        if (descents)
        {
            uint64_t down = ballot(i + 1 < n && key > keys[i + 1]);
            if (down && lane == __builtin_ctzll(act)) atomicAdd(&shDescents, uint32_t(__popcll(down)));
        }
        int first = __builtin_ctzll(act);
        for (int p = 0; p < numPasses; ++p)
        {
            // the digit of pass p as the pass kernel reads it: bits [shift, min(shift + 8, endBit))
            const int shift = beginBit + p * kRadixBits;
            const int bits  = endBit - shift < kRadixBits ? endBit - shift : kRadixBits;
            unsigned d      = unsigned(key >> shift) & ((1u << bits) - 1u);
            unsigned d0     = unsigned(__builtin_amdgcn_readlane(int(d), first));
            uint64_t same   = ballot(valid && d == d0);
            if (same == act)
            {
                // whole wave shares the digit (typical for high digits of nearly sorted keys)
                if (lane == first) atomicAdd(&sh[p * kRadix + d], uint32_t(__popcll(act)));
            }
            else if (valid) { atomicAdd(&sh[p * kRadix + d], 1u); }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < numPasses * kRadix; i += kThreads)
    {
        if (sh[i]) atomicAdd(&hist[i], sh[i]);
    }
    if (descents && threadIdx.x == 0 && shDescents) atomicAdd(descents, shDescents);
}

// x,y,z,h records of the particles in input order, written only when the input is far from key order (the
// descent count is final: this kernel runs after the histogram kernel on the same stream).
template<class T>
__global__ __launch_bounds__(256) void packRecordsKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                         const T* __restrict__ z, const T* __restrict__ h, size_t n,
                                                         const uint32_t* __restrict__ descents, uint32_t threshold,
                                                         T* __restrict__ packed)
{
    if (*descents <= threshold) return;
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double4>::type;
    V4 r;
    r.x                               = x[i];
    r.y                               = y[i];
    r.z                               = z[i];
    r.w                               = h ? h[i] : T(0);
    reinterpret_cast<V4*>(packed)[i] = r;
}

// Exclusive scan of each pass's 256-bin histogram, in place. One block of kRadix threads.
__global__ __launch_bounds__(kRadix) void scanHistogramKernel(uint32_t* hist, int numPasses)
{
    __shared__ uint32_t sh[kRadix];
    for (int p = 0; p < numPasses; ++p)
    {
        uint32_t v = hist[p * kRadix + threadIdx.x];
        sh[threadIdx.x] = v;
        __syncthreads();
        // Hillis-Steele inclusive scan over 256 values
        for (int off = 1; off < kRadix; off <<= 1)
        {
            uint32_t t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0u;
            __syncthreads();
            sh[threadIdx.x] += t;
            __syncthreads();
        }
        hist[p * kRadix + threadIdx.x] = sh[threadIdx.x] - v;
        __syncthreads();
    }
}

template<class T>
struct GatherArgs
{
    const T* xIn;
    const T* yIn;
    const T* zIn;
    const T* hIn;
    T* xOut;
    T* yOut;
    T* zOut;
    T* hOut;
    const T* packed;           // x,y,z,h records, used when *descents > threshold
    const uint32_t* descents;
    uint32_t threshold;
};

// ValueMode: 0 = keys only, 1 = values read from valuesIn, 2 = values are the input positions (iota)
// Gather: if GT is not void-like (GatherOn), the pass also writes xOut[dst] = xIn[value] etc.
template<class K, int ValueMode, bool GatherOn, class GT>
__global__ __launch_bounds__(kThreads) void onesweepPassKernel(const K* __restrict__ keysIn, K* __restrict__ keysOut,
                                                               const uint32_t* __restrict__ valuesIn,
                                                               uint32_t* __restrict__ valuesOut, size_t n, int shift,
                                                               int digitBits, const uint32_t* __restrict__ digitOffsets,
                                                               uint32_t* __restrict__ tileCounter,
                                                               uint32_t* __restrict__ status, GatherArgs<GT> g,
                                                               uint32_t* __restrict__ faults, uint32_t spinLimit)
{
    // the tile is reordered through LDS twice, keys first, then values in the same storage
    __shared__ union
    {
        K keys[kTile];
        uint32_t vals[kTile];
    } shTile;
    __shared__ uint32_t shWaveHist[kWaves][kRadix];
    __shared__ uint32_t shLocalStart[kRadix];
    __shared__ uint32_t shGlobalBase[kRadix];
    __shared__ uint32_t shTileId;
    __shared__ uint32_t shWaveSums[kWaves];

    const int tid   = threadIdx.x;
    const int lane  = laneId();
    const int wave  = tid / kWave;
    const uint32_t digitMask = (1u << digitBits) - 1u;

    if (tid == 0) shTileId = atomicAdd(tileCounter, 1u);
    for (int i = tid; i < kWaves * kRadix; i += kThreads)
        (&shWaveHist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = shTileId;
    const size_t tileBase = size_t(tile) * kTile;

    // ---- load (striped per wave: item k of lane l at waveBase + k*64 + l) ----
    K keys[kItems];
    uint32_t vals[kItems];
    uint32_t ranks[kItems];
    const size_t waveBase = tileBase + size_t(wave) * kWave * kItems;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i = waveBase + size_t(k) * kWave + lane;
        bool v   = i < n;
        keys[k]  = v ? keysIn[i] : K(0);
        if constexpr (ValueMode == 1) { vals[k] = v ? valuesIn[i] : 0u; }
        else if constexpr (ValueMode == 2) { vals[k] = uint32_t(i); }
    }

    // ---- stable rank within the wave: match lanes with equal digits by ballots ----
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i       = waveBase + size_t(k) * kWave + lane;
        bool v         = i < n;
        unsigned d     = unsigned(keys[k] >> shift) & digitMask;
        uint64_t peers = ballot(v);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b)
        {
            if (b < digitBits)
            {
                bool bit     = (d >> b) & 1u;
                uint64_t bal = ballot(bit);
                peers &= bit ? bal : ~bal;
            }
        }
        unsigned before = laneRank(peers);
        uint32_t prev   = v ? shWaveHist[wave][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (v && before == 0) shWaveHist[wave][d] = prev + uint32_t(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
        ranks[k] = prev + before;
    }
    __syncthreads();

    // ---- per digit: counts of the tile, exclusive offsets of each wave ----
    uint32_t tileCount = 0;
    {
        const int d = tid; // kThreads == kRadix
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w)
        {
            uint32_t c       = shWaveHist[w][d];
            shWaveHist[w][d] = run;
            run += c;
        }
        tileCount = run;
    }

    // publish this tile's counts for the look-back before scanning locally
    uint32_t* myStatus = status + size_t(tile) * kRadix;
    if (tile == 0) atomicStoreRelaxed(myStatus + tid, kFlagPrefix | tileCount);
    else atomicStoreRelaxed(myStatus + tid, kFlagAggregate | tileCount);

    // ---- block exclusive scan of tileCount over the 256 digits ----
    uint32_t incl = tileCount;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1)
    {
        uint32_t t = uint32_t(__shfl_up(int(incl), off, kWave));
        if (lane >= off) incl += t;
    }
    if (lane == kWave - 1) shWaveSums[wave] = incl;

    // ---- decoupled look-back: exclusive prefix of this digit over all earlier tiles ----
    uint32_t exclusive = 0;
    if (tile > 0)
    {
        // status words of other tiles live in other XCDs' caches, so each read is a long round trip: read a window
        // of kLookBack predecessors per trip and consume them nearest first, up to the first unpublished one
        constexpr int kLookBack = 4;
        int64_t j = int64_t(tile) - 1;
        uint32_t spins = 0;
        bool done = false;
        while (!done && j >= 0)
        {
            uint32_t w[kLookBack];
#pragma unroll
            for (int q = 0; q < kLookBack; ++q)
                w[q] = j - q >= 0 ? atomicLoadRelaxed(status + size_t(j - q) * kRadix + tid) : kFlagPrefix;
            bool stalled = false;
#pragma unroll
            for (int q = 0; q < kLookBack; ++q)
            {
                if (done || stalled) continue;
                uint32_t flag = w[q] & ~kValueMask;
                if (flag == 0)
                {
                    stalled = true;
                    continue;
                }
                exclusive += w[q] & kValueMask;
                if (flag == kFlagPrefix) done = true;
                else --j;
            }
            if (stalled)
            {
                // a predecessor has not published yet; every tile publishes before it waits, so this ends. The
                // cap only guards against a corrupted workspace: giving up leaves a wrong prefix, which the fault
                // word reports
                if (++spins > spinLimit)
                {
                    atomicOr(faults, HIPCSTONE_FAULT_SORT_LOOKBACK);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        atomicStoreRelaxed(myStatus + tid, kFlagPrefix | (exclusive + tileCount));
    }
    __syncthreads();
    uint32_t waveOffset = 0;
    for (int w = 0; w < wave; ++w)
        waveOffset += shWaveSums[w];
    uint32_t localStart   = waveOffset + incl - tileCount;
    shLocalStart[tid]     = localStart;
    shGlobalBase[tid]     = digitOffsets[tid] + exclusive - localStart;
    __syncthreads();

    // ---- scatter the keys into LDS in sorted order (ranks[k] becomes the key's slot in the tile) ----
    uint32_t validInTile = n > tileBase ? uint32_t(min(size_t(kTile), n - tileBase)) : 0u;
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i = waveBase + size_t(k) * kWave + lane;
        if (i < n)
        {
            unsigned d        = unsigned(keys[k] >> shift) & digitMask;
            ranks[k]          = shLocalStart[d] + shWaveHist[wave][d] + ranks[k];
            shTile.keys[ranks[k]] = keys[k];
        }
    }
    __syncthreads();

    // ---- write out: consecutive LDS slots of one digit go to consecutive global slots; slot i = tid + m * 256 ----
    uint32_t dsts[kItems];
#pragma unroll
    for (int m = 0; m < kItems; ++m)
    {
        uint32_t i = uint32_t(tid) + uint32_t(m) * kThreads;
        if (i < validInTile)
        {
            K key        = shTile.keys[i];
            unsigned d   = unsigned(key >> shift) & digitMask;
            dsts[m]      = shGlobalBase[d] + i; // modulo 2^32: the base alone may wrap below zero
            // a correct prefix always lands inside [0, n); after a look-back give-up (reported through the fault
            // word) later passes see digit counts that disagree with the histogram, and this keeps them in bounds
            if (dsts[m] < n) keysOut[dsts[m]] = key;
        }
    }
    if constexpr (ValueMode == 0) return;
    __syncthreads(); // every key is read out before the values reuse the storage
#pragma unroll
    for (int k = 0; k < kItems; ++k)
    {
        size_t i = waveBase + size_t(k) * kWave + lane;
        if (i < n) shTile.vals[ranks[k]] = vals[k];
    }
    __syncthreads();

    bool usePacked = false;
    if constexpr (GatherOn) usePacked = g.packed && *g.descents > g.threshold;
#pragma unroll
    for (int m = 0; m < kItems; ++m)
    {
        uint32_t i = uint32_t(tid) + uint32_t(m) * kThreads;
        if (i < validInTile)
        {
            const uint32_t dst = dsts[m];
            uint32_t v         = shTile.vals[i];
            if (dst >= n) continue; // only after a look-back give-up (see above)
            valuesOut[dst]     = v;
            if constexpr (GatherOn)
            {
                if (v >= n) continue; // gather values are input positions; see above
                if (usePacked)
                {
                    using V4 = typename std::conditional<sizeof(GT) == 4, float4, double4>::type;
                    V4 r     = reinterpret_cast<const V4*>(g.packed)[v];
                    g.xOut[dst] = r.x;
                    g.yOut[dst] = r.y;
                    g.zOut[dst] = r.z;
                    if (g.hIn) g.hOut[dst] = r.w;
                }
                else
                {
                    g.xOut[dst] = g.xIn[v];
                    g.yOut[dst] = g.yIn[v];
                    g.zOut[dst] = g.zIn[v];
                    if (g.hIn) g.hOut[dst] = g.hIn[v];
                }
            }
        }
    }
}

template<class T>
__global__ void gatherKernel(const uint32_t* __restrict__ perm, size_t n, const T* __restrict__ src,
                             T* __restrict__ dst)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ void fillU32Kernel(uint32_t* p, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static hipcstone_status zeroU32(uint32_t* p, size_t n, hipStream_t s)
{
    if (n == 0) return HIPCSTONE_OK;
    fillU32Kernel<<<ceilDiv(n, 256), 256, 0, s>>>(p, n);
    HC_CHECK_LAUNCH("fillU32Kernel");
    return HIPCSTONE_OK;
}

template<class K>
static int numPassesFor(int endBit)
{
    return (endBit + kRadixBits - 1) / kRadixBits;
}

// Generic driver. valueMode 0/1/2 as in onesweepPassKernel. When gather is set, the last pass gathers.
template<class K, class GT>
static hipcstone_status radixSort(hipStream_t s, K* keys, uint32_t* values, size_t n, K* keyBuf, uint32_t* valueBuf,
                                  void* workspace, size_t wsBytes, int endBit, int valueMode, const GatherArgs<GT>* g,
                                  int beginBit = 0)
{
    if (n == 0) return HIPCSTONE_OK;
    if (n > kValueMask) return invalid("radix sort: more than 2^30-1 keys");
    if (endBit <= 0 || endBit > int(8 * sizeof(K))) endBit = int(8 * sizeof(K));
    if (beginBit < 0 || beginBit >= endBit) return invalid("radix sort: begin bit outside [0, end bit)");
    int passes = numPassesFor<K>(endBit - beginBit);
    const bool pack = g != nullptr;
    size_t need     = workspaceBytes(n, passes, pack);
    if (!workspace || wsBytes < need) return invalid("radix sort: workspace too small");
    if (!keys || !keyBuf) return invalid("radix sort: null key buffer");
    if (valueMode != 0 && (!values || !valueBuf)) return invalid("radix sort: null value buffer");

    SortWorkspace w = carveWorkspace(workspace, n, passes, pack);
    size_t tiles    = numTiles(n);
    uint32_t* faults         = faultWord();
    const uint32_t spinLimit = sortLookbackLimit();
    if (!faults) return HIPCSTONE_HIP_ERROR;
    if (auto st = zeroU32(w.hist, size_t(kMaxPasses) * kRadix, s)) return st;
    if (auto st = zeroU32(w.counter, kMaxPasses + 1, s)) return st;
    if (auto st = zeroU32(w.status, size_t(passes) * tiles * kRadix, s)) return st;

    unsigned hb = unsigned(std::min<size_t>(kHistBlocks, (n + kThreads - 1) / kThreads));
    uint32_t* descents = pack ? w.counter + kMaxPasses : nullptr;
    histogramKernel<K><<<hb, kThreads, 0, s>>>(keys, n, passes, beginBit, endBit, w.hist, descents);
    HC_CHECK_LAUNCH("histogramKernel");
    GatherArgs<GT> gp{};
    if (pack)
    {
        gp           = *g;
        gp.packed    = static_cast<const GT*>(w.packed);
        gp.descents  = descents;
        gp.threshold = uint32_t(n / 8);
        packRecordsKernel<GT><<<ceilDiv(n, 256), 256, 0, s>>>(g->xIn, g->yIn, g->zIn, g->hIn, n, descents,
                                                             gp.threshold, static_cast<GT*>(w.packed));
        HC_CHECK_LAUNCH("packRecordsKernel");
    }
    scanHistogramKernel<<<1, kRadix, 0, s>>>(w.hist, passes);
    HC_CHECK_LAUNCH("scanHistogramKernel");

    // ping-pong so that the last pass writes into keys/values
    K* src         = keys;
    uint32_t* vsrc = values;
    K* dst;
    uint32_t* vdst;
    GatherArgs<GT> noGather{};
    for (int p = 0; p < passes; ++p)
    {
        bool last = p == passes - 1;
        // destination alternates; arrange the final destination to be keys
        bool toBuf = ((passes - 1 - p) % 2) == 1;
        dst        = toBuf ? keyBuf : keys;
        vdst       = toBuf ? valueBuf : values;
        int shift  = beginBit + p * kRadixBits;
        int bits   = std::min(kRadixBits, endBit - shift);
        int mode   = (p == 0 && valueMode == 2) ? 2 : (valueMode == 0 ? 0 : 1);
        const uint32_t* off = w.hist + p * kRadix;
        uint32_t* cnt       = w.counter + p;
        uint32_t* st        = w.status + size_t(p) * tiles * kRadix;
        // an odd number of passes would make pass 0 read and write keys; stage through the buffer first
        if (src == dst)
        {
            HC_CHECK(hipMemcpyAsync(keyBuf, keys, n * sizeof(K), hipMemcpyDeviceToDevice, s));
            src = keyBuf;
            if (mode == 1)
            {
                HC_CHECK(hipMemcpyAsync(valueBuf, values, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
                vsrc = valueBuf;
            }
        }
        bool gatherNow = last && g != nullptr;
        if (mode == 0)
            onesweepPassKernel<K, 0, false, GT><<<unsigned(tiles), kThreads, 0, s>>>(
                src, dst, nullptr, nullptr, n, shift, bits, off, cnt, st, noGather, faults, spinLimit);
        else if (mode == 2 && !gatherNow)
            onesweepPassKernel<K, 2, false, GT><<<unsigned(tiles), kThreads, 0, s>>>(
                src, dst, nullptr, vdst, n, shift, bits, off, cnt, st, noGather, faults, spinLimit);
        else if (mode == 2 && gatherNow)
            onesweepPassKernel<K, 2, true, GT><<<unsigned(tiles), kThreads, 0, s>>>(
                src, dst, nullptr, vdst, n, shift, bits, off, cnt, st, gp, faults, spinLimit);
        else if (!gatherNow)
            onesweepPassKernel<K, 1, false, GT><<<unsigned(tiles), kThreads, 0, s>>>(
                src, dst, vsrc, vdst, n, shift, bits, off, cnt, st, noGather, faults, spinLimit);
        else
            onesweepPassKernel<K, 1, true, GT><<<unsigned(tiles), kThreads, 0, s>>>(
                src, dst, vsrc, vdst, n, shift, bits, off, cnt, st, gp, faults, spinLimit);
        HC_CHECK_LAUNCH("onesweepPassKernel");
        src  = dst;
        vsrc = vdst;
    }
    return HIPCSTONE_OK;
}

template<class K>
static size_t sortWorkspaceSize(size_t n, bool packed)
{
    return workspaceBytes(n, numPassesFor<K>(int(8 * sizeof(K))), packed);
}

template<class T>
static hipcstone_status gather(hipcstone_stream s, const uint32_t* perm, size_t n, const T* src, T* dst)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!perm || !src || !dst) return invalid("gather: null pointer");
    gatherKernel<T><<<ceilDiv(n, 256), 256, 0, stream(s)>>>(perm, n, src, dst);
    HC_CHECK_LAUNCH("gatherKernel");
    return HIPCSTONE_OK;
}

template<class K>
size_t sortGatherWorkspaceBytes(size_t n)
{
    return sortWorkspaceSize<K>(n, true);
}

template<class K, class T>
hipcstone_status sortGatherBits(hipStream_t s, K* keys, uint32_t* perm, size_t n, const T* xIn, const T* yIn,
                                const T* zIn, const T* hIn, T* xOut, T* yOut, T* zOut, T* hOut, K* keyBuf,
                                uint32_t* valueBuf, void* ws, size_t wsBytes, int beginBit, int endBit)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!xIn || !yIn || !zIn || !xOut || !yOut || !zOut || (hIn && !hOut)) return invalid("sortGather: null array");
    GatherArgs<T> g{xIn, yIn, zIn, hIn, xOut, yOut, zOut, hOut, nullptr, nullptr, 0};
    return radixSort<K, T>(s, keys, perm, n, keyBuf, valueBuf, ws, wsBytes, endBit, 2, &g, beginBit);
}

template size_t sortGatherWorkspaceBytes<uint64_t>(size_t);
template hipcstone_status sortGatherBits<uint64_t, float>(hipStream_t, uint64_t*, uint32_t*, size_t, const float*,
                                                          const float*, const float*, const float*, float*, float*,
                                                          float*, float*, uint64_t*, uint32_t*, void*, size_t, int,
                                                          int);

template<class K, class T>
static hipcstone_status sortGather(hipcstone_stream s, K* keys, uint32_t* perm, size_t n, const T* xIn, const T* yIn,
                                   const T* zIn, const T* hIn, T* xOut, T* yOut, T* zOut, T* hOut, K* keyBuf,
                                   uint32_t* valueBuf, void* ws, size_t wsBytes)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!xIn || !yIn || !zIn || !xOut || !yOut || !zOut || (hIn && !hOut)) return invalid("sortGather: null array");
    GatherArgs<T> g{xIn, yIn, zIn, hIn, xOut, yOut, zOut, hOut, nullptr, nullptr, 0};
    return radixSort<K, T>(stream(s), keys, perm, n, keyBuf, valueBuf, ws, wsBytes, int(8 * sizeof(K)), 2, &g);
}

// ---- the no-allocation sort: its workspace lives in a file-scope device array of fixed size, sized for the
// largest accepted key count of 64-bit keys with all 8 passes. Each device holds its own copy of the array; host
// code never uses the array's name as an address, only the per-device pointer hipGetSymbolAddress returned.
constexpr size_t kNoAllocMaxKeys = size_t(1) << 24;
constexpr size_t kNoAllocScratchBytes =
    ((sizeof(uint32_t) * kMaxPasses * kRadix + 255) & ~size_t(255)) + ((sizeof(uint32_t) * (kMaxPasses + 1) + 255) &
                                                                        ~size_t(255)) +
    ((sizeof(uint32_t) * size_t(kMaxPasses) * ((kNoAllocMaxKeys + kTile - 1) / kTile) * kRadix + 255) & ~size_t(255));

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratchMutex;
static void* g_scratchPtr[kNoAllocMaxDevices];

// the scratch array's address on the current device, looked up once per device with that device current
static void* noAllocScratch()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices)
    {
        setError("sort (no allocation): no current device");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev])
    {
        void* ptr    = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
        {
            setError(std::string("sort (no allocation): hipGetSymbolAddress failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratchPtr[dev] = ptr;
    }
    return g_scratchPtr[dev];
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

hipcstone_status hipcstone_sort_by_key_u64_noalloc(hipcstone_stream s, uint64_t* keys, uint32_t* values, size_t n,
                                                   uint64_t* key_buf, uint32_t* value_buf, int end_bit)
{
    if (n > kNoAllocMaxKeys)
        return invalid("sort_by_key_u64_noalloc: more than 2^24 keys (the static scratch's maximum)");
    void* ptr = noAllocScratch();
    if (!ptr) return HIPCSTONE_HIP_ERROR;
    static_assert(kNoAllocScratchBytes >= 8192, "scratch holds at least the histograms");
    // every part of the workspace that must start at zero is cleared by radixSort itself (fill kernels)
    return radixSort<uint64_t, float>(stream(s), keys, values, n, key_buf, value_buf, ptr, kNoAllocScratchBytes,
                                      end_bit, values ? 1 : 0, nullptr);
}

size_t hipcstone_sort_by_key_workspace_size_u32(size_t n) { return sortWorkspaceSize<uint32_t>(n, false); }
size_t hipcstone_sort_by_key_workspace_size_u64(size_t n) { return sortWorkspaceSize<uint64_t>(n, false); }
size_t hipcstone_sort_gather_workspace_size_u32(size_t n) { return sortWorkspaceSize<uint32_t>(n, true); }
size_t hipcstone_sort_gather_workspace_size_u64(size_t n) { return sortWorkspaceSize<uint64_t>(n, true); }

hipcstone_status hipcstone_sort_by_key_u32(hipcstone_stream s, uint32_t* keys, uint32_t* values, size_t n,
                                           uint32_t* key_buf, uint32_t* value_buf, void* workspace,
                                           size_t workspace_bytes, int end_bit)
{
    return radixSort<uint32_t, float>(stream(s), keys, values, n, key_buf, value_buf, workspace, workspace_bytes,
                                      end_bit, values ? 1 : 0, nullptr);
}

hipcstone_status hipcstone_sort_by_key_u64(hipcstone_stream s, uint64_t* keys, uint32_t* values, size_t n,
                                           uint64_t* key_buf, uint32_t* value_buf, void* workspace,
                                           size_t workspace_bytes, int end_bit)
{
    return radixSort<uint64_t, float>(stream(s), keys, values, n, key_buf, value_buf, workspace, workspace_bytes,
                                      end_bit, values ? 1 : 0, nullptr);
}

#define HIPCSTONE_DEF_SORTGATHER(KS, K, RS, T)                                                                         \
    hipcstone_status hipcstone_sort_gather_##KS##_##RS(                                                                \
        hipcstone_stream s, K* keys, uint32_t* perm, size_t n, const T* x_in, const T* y_in, const T* z_in,           \
        const T* h_in, T* x_out, T* y_out, T* z_out, T* h_out, K* key_buf, uint32_t* value_buf, void* workspace,       \
        size_t workspace_bytes)                                                                                        \
    {                                                                                                                  \
        return sortGather<K, T>(s, keys, perm, n, x_in, y_in, z_in, h_in, x_out, y_out, z_out, h_out, key_buf,        \
                                value_buf, workspace, workspace_bytes);                                                \
    }

HIPCSTONE_DEF_SORTGATHER(u32, uint32_t, f32, float)
HIPCSTONE_DEF_SORTGATHER(u32, uint32_t, f64, double)
HIPCSTONE_DEF_SORTGATHER(u64, uint64_t, f32, float)
HIPCSTONE_DEF_SORTGATHER(u64, uint64_t, f64, double)

hipcstone_status hipcstone_gather_u32(hipcstone_stream s, const uint32_t* perm, size_t n, const uint32_t* src,
                                      uint32_t* dst)
{
    return gather<uint32_t>(s, perm, n, src, dst);
}

hipcstone_status hipcstone_gather_u64(hipcstone_stream s, const uint32_t* perm, size_t n, const uint64_t* src,
                                      uint64_t* dst)
{
    return gather<uint64_t>(s, perm, n, src, dst);
}

} // extern "C"
