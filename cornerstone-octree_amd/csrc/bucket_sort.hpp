// Stable sort + gather of particles by SFC key through the buckets of a previous leaf array (warm steps of the
// pipeline and of the one-rank Domain). Included by the translation units that use it.
#pragma once

#include <algorithm>

#include "common.hpp"
#include "host_util.hpp"
#include "runtime.hpp"

namespace hipcstone
{

// rank += ((kq, iq) < (k, id)): the borrow out of the multi-word subtraction (kq:iq) - (k:id), added with carry
__device__ inline uint32_t addIfPairLess(uint32_t rank, uint64_t kq, uint32_t iq, uint64_t k, uint32_t id)
{
    uint32_t t;
    asm("v_sub_co_u32_e32 %[t], vcc, %[iq], %[id]\n\t"
        "v_subb_co_u32_e32 %[t], vcc, %[ql], %[kl], vcc\n\t"
        "v_subb_co_u32_e32 %[t], vcc, %[qh], %[kh], vcc\n\t"
        "v_addc_co_u32_e32 %[r], vcc, 0, %[r], vcc"
        : [t] "=&v"(t), [r] "+v"(rank)
        : [iq] "v"(iq), [id] "v"(id), [ql] "v"(uint32_t(kq)), [kl] "v"(uint32_t(k)), [qh] "v"(uint32_t(kq >> 32)),
          [kh] "v"(uint32_t(k >> 32))
        : "vcc");
    return rank;
}

__device__ inline uint32_t addIfPairLess(uint32_t rank, uint32_t kq, uint32_t iq, uint32_t k, uint32_t id)
{
    uint32_t t;
    asm("v_sub_co_u32_e32 %[t], vcc, %[iq], %[id]\n\t"
        "v_subb_co_u32_e32 %[t], vcc, %[kq], %[k], vcc\n\t"
        "v_addc_co_u32_e32 %[r], vcc, 0, %[r], vcc"
        : [t] "=&v"(t), [r] "+v"(rank)
        : [iq] "v"(iq), [id] "v"(id), [kq] "v"(kq), [k] "v"(k)
        : "vcc");
    return rank;
}

// ---- warm-step sort by the previous leaf array.
// The stable sort of keys with values iota (gather.hpp:42-67, primitives_gpu.cu:310-367) is the unique ordering by
// (key, input index). Any partition of the key space into ordered ranges gives it as the concatenation of each
// range's elements ordered by (key, index); the previous step's leaf array is such a partition, and after a small
// displacement each of its leaves holds about bucketSize keys. So a warm step buckets every key into the previous
// leaf that contains it (counted with atomics, in any order), scatters the buckets contiguously, and orders each bucket
// by (key, index) in one wave, gathering x,y,z,h on the way. The result equals the radix sort's for any input: the
// previous tree only decides how the work is split.

template<class K>
__device__ inline K waveMinKey(K v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
    {
        K o;
        if constexpr (sizeof(K) == 8)
        {
            uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), off, kWave));
            uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), off, kWave));
            o           = K((uint64_t(hi) << 32) | lo);
        }
        else { o = K(uint32_t(__shfl_xor(int(v), off, kWave))); }
        v = o < v ? o : v;
    }
    return v;
}

template<class K>
__device__ inline K waveMaxKey(K v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
    {
        K o;
        if constexpr (sizeof(K) == 8)
        {
            uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(v)), off, kWave));
            uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(v >> 32)), off, kWave));
            o           = K((uint64_t(hi) << 32) | lo);
        }
        else { o = K(uint32_t(__shfl_xor(int(v), off, kWave))); }
        v = o > v ? o : v;
    }
    return v;
}

// one round of a 64-ary search for the wave-uniform key k with leaves[lo] <= k < leaves[hi]: pivot of this lane
// (span <= 64: lo + lane, exact in one round; else 64 strictly increasing pivots with pivot 0 = lo)
__device__ inline uint32_t searchPivot(uint32_t lo, uint32_t hi, int lane, bool& on)
{
    const uint32_t span = hi - lo;
    on                  = span > 1 && (span <= uint32_t(kWave) ? uint32_t(lane) < span : true);
    const uint32_t piv  = span <= uint32_t(kWave) ? lo + uint32_t(lane) : lo + uint32_t((uint64_t(span) * uint32_t(lane)) >> 6);
    return on ? piv : lo;
}

// update of the bracket from the ballot of (pivot <= k): lanes 0..c-1, c >= 1 (pivot 0 is lo)
__device__ inline void searchNarrow(uint32_t& lo, uint32_t& hi, uint32_t piv, uint64_t le)
{
    if (hi - lo <= 1) return;
    const uint32_t span = hi - lo;
    const int c         = __popcll(le);
    const uint32_t nlo  = uint32_t(__builtin_amdgcn_readlane(int(piv), c - 1));
    if (span <= uint32_t(kWave)) hi = nlo + 1;
    else if (c < kWave) hi = uint32_t(__builtin_amdgcn_readlane(int(piv), c));
    lo = nlo;
}

// the leaves holding the wave-uniform keys kmin <= kmax: both 64-ary searches advance in the same rounds, so their
// pivot loads overlap
template<class K>
__device__ inline void waveLeafRange(const K* __restrict__ leaves, uint32_t L, K kmin, K kmax, int lane, uint32_t& a,
                                     uint32_t& b)
{
    uint32_t lo0 = 0, hi0 = L, lo1 = 0, hi1 = L;
    while (hi0 - lo0 > 1 || hi1 - lo1 > 1)
    {
        bool on0, on1;
        const uint32_t p0 = searchPivot(lo0, hi0, lane, on0);
        const uint32_t p1 = searchPivot(lo1, hi1, lane, on1);
        const K v0        = leaves[p0];
        const K v1        = leaves[p1];
        const uint64_t l0 = ballot(on0 && v0 <= kmin);
        const uint64_t l1 = ballot(on1 && v1 <= kmax);
        searchNarrow(lo0, hi0, p0, l0);
        searchNarrow(lo1, hi1, p1, l1);
    }
    a = lo0;
    b = lo1;
}

// Leaves the hint looks through: the wave's keys are searched among leaves [hint, hint + kHintLeaves)
constexpr uint32_t kHintLeaves = 16;

// hint[g] = the leaf of the previous layout that holds position 64 g (the previous step's leaf of the particle that
// now sits first in wave g): one thread per leaf writes the waves whose first position falls in it. layout[0..L] is
// the previous leaf layout of exactly n particles.
template<int Unused = 0>
__global__ __launch_bounds__(256) void waveLeafHintKernel(const uint32_t* __restrict__ layout, int32_t L, uint32_t n,
                                                          uint32_t* __restrict__ hint)
{
    const uint32_t l = blockIdx.x * 256u + threadIdx.x;
    if (l >= uint32_t(L)) return;
    const uint32_t a = layout[l], b = layout[l + 1];
    for (uint32_t g = (a + 63u) / 64u; 64u * g < b && 64u * g < n; ++g)
        hint[g] = l;
}

// bucket[i] = the leaf of the previous leaf array holding keys[i]; slot[i] = its position among its bucket's keys
// (any order). segCount[leaf] counts them; it starts at zero. hint (optional, from waveLeafHintKernel): when every
// key of a wave lies in the hint's kHintLeaves leaves, the wave counts boundaries there and skips the search; any
// other wave searches the whole leaf array (the hint only saves work, it never changes a bucket).
template<class K>
__global__ __launch_bounds__(256) void bucketByLeafKernel(const K* __restrict__ keys, uint32_t n,
                                                          const K* __restrict__ leaves, int32_t L,
                                                          uint32_t* __restrict__ segCount, uint32_t* __restrict__ bucket,
                                                          uint32_t* __restrict__ slot,
                                                          const uint32_t* __restrict__ hint)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const int lane   = laneId();
    const bool valid = i < n;
    if (ballot(valid) == 0) return; // wave-uniform
    // every lane loads (a clamped index past the end), so the key and the hint are in flight together
    const K key = keys[valid ? i : n - 1];
    uint32_t lo = 0;
    bool found  = false;
    if (hint)
    {
        const uint32_t l0 = uniform(hint[uniform(i / 64u)]);
        if (uint32_t(L) >= kHintLeaves && l0 <= uint32_t(L) - kHintLeaves) // no wrap for any hint value
        {
            // the hint's kHintLeaves + 1 boundaries with wave-uniform addresses: a few wide scalar loads issued
            // together instead of one dependent load per boundary
            const K* __restrict__ window = leaves + l0;
            K lv[kHintLeaves + 1];
#pragma unroll
            for (uint32_t t = 0; t <= kHintLeaves; ++t)
                lv[t] = window[t];
            // counted before the window test, so that all the window's loads are issued before any is waited for
            uint32_t c = l0;
#pragma unroll
            for (uint32_t t = 1; t < kHintLeaves; ++t)
                c += uint32_t(lv[t] <= key);
            const bool inside = !valid || (lv[0] <= key && key < lv[kHintLeaves]);
            if (ballot(!inside) == 0)
            {
                lo    = c;
                found = true;
            }
        }
    }
    if (!found)
    {
        // the leaf range holding the wave's keys, then a binary search inside it per lane
        // (keys lie in [leaves[0], leaves[L]) = [0, nodeRange(0)))
        const K kmin = waveMinKey(valid ? key : ~K(0));
        const K kmax = waveMaxKey(valid ? key : K(0));
        uint32_t a, b;
        waveLeafRange(leaves, uint32_t(L), kmin, kmax, lane, a, b);
        lo = a; // leaves[a] <= key < leaves[b + 1]
        if (b - a <= 16)
        {
            // few leaves: count the boundaries at or below the key (wave-uniform addresses)
            for (uint32_t j = a + 1; j <= b; ++j)
                lo += uint32_t(leaves[j] <= key);
        }
        else
        {
            uint32_t hi = b + 1;
            while (hi - lo > 1)
            {
                const uint32_t mid = (lo + hi) >> 1;
                if (leaves[mid] <= key) lo = mid;
                else hi = mid;
            }
        }
    }
    // the wave's lanes grouped by leaf: the first lane of each group counts it, then one atomic instruction, issued
    // by the first lanes of all groups together, adds every group's size (one round trip per wave, not per group)
    uint64_t todo    = ballot(valid);
    uint32_t myCount = 0; // the group's size, in its first lane
    uint32_t myRank  = 0; // rank among the group's lanes
    int myLeader     = 0;
    while (todo)
    {
        const int leader     = __ffsll((unsigned long long)todo) - 1;
        const uint32_t g     = uint32_t(__builtin_amdgcn_readlane(int(lo), leader));
        const uint64_t peers = ballot(valid && lo == g);
        if (lane == leader) myCount = uint32_t(__popcll(peers));
        if (laneIn(peers))
        {
            myLeader = leader;
            myRank   = laneRank(peers);
        }
        todo &= ~peers;
    }
    if (valid) bucket[i] = lo;
    uint32_t myBase = 0; // the atomic's result, in the first lane of each group
    if (myCount != 0) myBase = atomicAdd(&segCount[lo], myCount);
    const uint32_t base = uint32_t(__shfl(int(myBase), myLeader, kWave));
    if (valid) slot[i] = base + myRank;
}

// largest bucket: each block reduces kMaxPerBlock counts, whose loads are all in flight together, and makes at most
// one atomicMax into *out (zeroed by the caller); few blocks keep the atomics on that one word few
constexpr uint32_t kMaxPerThread = 16;
constexpr uint32_t kMaxPerBlock  = 256u * kMaxPerThread;

template<int Unused = 0>
__global__ __launch_bounds__(256) void maxBucketKernel(const uint32_t* __restrict__ counts, uint32_t n,
                                                       uint32_t* __restrict__ out)
{
    __shared__ uint32_t sh[4];
    const uint64_t base = uint64_t(blockIdx.x) * kMaxPerBlock + threadIdx.x;
    uint32_t v[kMaxPerThread];
#pragma unroll
    for (uint32_t k = 0; k < kMaxPerThread; ++k)
    {
        const uint64_t i = base + uint64_t(k) * 256u;
        v[k]             = i < n ? counts[i] : 0u;
    }
    uint32_t m = 0;
#pragma unroll
    for (uint32_t k = 0; k < kMaxPerThread; ++k)
        m = v[k] > m ? v[k] : m;
    for (int off = 32; off >= 1; off >>= 1)
    {
        const uint32_t o = uint32_t(__shfl_xor(int(m), off, kWave));
        m                = o > m ? o : m;
    }
    if (laneId() == 0) sh[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const uint32_t b = std::max(std::max(sh[0], sh[1]), std::max(sh[2], sh[3]));
        if (b > *out) atomicMax(out, b); // the pre-check skips most atomics
    }
}

template<int Unused = 0>
__global__ void zeroU32Kernel(uint32_t* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) p[i] = 0u;
}

template<class K>
__global__ __launch_bounds__(256) void scatterByBucketKernel(const K* __restrict__ keys, uint32_t n,
                                                             const uint32_t* __restrict__ bucket,
                                                             const uint32_t* __restrict__ slot,
                                                             const uint32_t* __restrict__ segOffset,
                                                             K* __restrict__ keysOut, uint32_t* __restrict__ idxOut)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t pos = segOffset[bucket[i]] + slot[i];
    keysOut[pos]       = keys[i];
    idxOut[pos]        = i;
}

// Each wave orders the buckets [g * kBucketsPerWave, (g + 1) * kBucketsPerWave), which are contiguous in the scattered
// arrays, 64 elements at a time: a window starting at a bucket boundary holds some whole buckets (and the start of
// the next one, left for the next window); every lane of a whole bucket counts the elements of its bucket that come
// before it in (key, index) order (the pairs are distinct) and writes its element, with x,y,z,h, to bucket start +
// rank. A bucket of more than 64 elements is ordered by itself, its elements compared in chunks of 64. Each element's
// bucket range comes from the wave's 17 bucket offsets, held in scalar registers (no per-element range in memory).
constexpr uint32_t kBucketsPerWave = 16;

template<class K, class T>
__global__ __launch_bounds__(256) void sortBucketsGatherKernel(const K* __restrict__ keysIn,
                                                               const uint32_t* __restrict__ idxIn,
                                                               const uint32_t* __restrict__ segOffset, int32_t L,
                                                               K* __restrict__ keysOut, uint32_t* __restrict__ permOut,
                                                               const T* __restrict__ x, const T* __restrict__ y,
                                                               const T* __restrict__ z, const T* __restrict__ h,
                                                               T* __restrict__ xo, T* __restrict__ yo,
                                                               T* __restrict__ zo, T* __restrict__ ho)
{
    __shared__ K shKey[256];
    __shared__ uint32_t shIdx[256];
    const int lane = laneId();
    const int w    = uniform(int(threadIdx.x) / kWave); // wave-uniform, so the offsets below are scalar loads
    K* wKey        = shKey + w * kWave;
    uint32_t* wIdx = shIdx + w * kWave;
    const uint64_t g = uint64_t(blockIdx.x) * 4 + uint64_t(w);
    if (g * kBucketsPerWave >= uint64_t(L)) return; // wave-uniform
    const uint32_t b0  = uint32_t(g * kBucketsPerWave);
    const uint32_t b1  = uint32_t(std::min<uint64_t>(uint64_t(b0) + kBucketsPerWave, uint64_t(L)));
    const uint32_t end = uniform(segOffset[b1]);
    // the offsets of the wave's buckets (offset k of bucket b0 + k; past b1 they repeat the end): for all but the
    // last wave, one window of kBucketsPerWave + 1 offsets (a few wide scalar loads)
    uint32_t off[kBucketsPerWave + 1];
    if (b1 == b0 + kBucketsPerWave)
    {
        const uint32_t* __restrict__ window = segOffset + b0;
#pragma unroll
        for (uint32_t k = 0; k <= kBucketsPerWave; ++k)
            off[k] = window[k];
    }
    else
    {
#pragma unroll
        for (uint32_t k = 0; k <= kBucketsPerWave; ++k)
            off[k] = uniform(segOffset[b0 + k < b1 ? b0 + k : b1]);
    }
    auto put = [&](uint32_t pos, K k, uint32_t id) {
        keysOut[pos] = k;
        permOut[pos] = id;
        xo[pos]      = x[id];
        yo[pos]      = y[id];
        zo[pos]      = z[id];
        ho[pos]      = h[id];
    };
    // A window starts at a bucket boundary P: each lane's element P + lane, its bucket [r.x, r.y) and input index.
    // The buckets come from the offsets alone, so the wave knows where every window starts before any key arrives,
    // and runs two windows ahead: the keys and indices of window 2 and the x,y,z,h of window 1 are in flight while
    // window 0 is ranked and stored.
    struct Window
    {
        uint32_t P;
        uint2 r;
        K k;
        uint32_t id;
    };
    auto open = [&](Window& W, uint32_t p) {
        W.P              = p;
        const uint32_t e = p + uint32_t(lane);
        const bool valid = e < end;
        // the bucket [off[j], off[j + 1]) holding e: the last offset at or below e and the first above it (the
        // offsets ascend; empty buckets repeat an offset)
        uint32_t rx = off[0], ry = ~0u;
#pragma unroll
        for (uint32_t j = 1; j <= kBucketsPerWave; ++j)
        {
            rx = off[j] <= e ? off[j] : rx;
            ry = off[j] > e && off[j] < ry ? off[j] : ry;
        }
        W.r  = valid ? make_uint2(rx, ry) : make_uint2(0u, 0u);
        W.k  = valid ? keysIn[e] : K(0);
        W.id = valid ? idxIn[e] : 0u;
    };
    // the end of lane 0's bucket: a window whose first bucket holds more than 64 elements is that bucket alone
    auto firstEnd = [&](const Window& W) -> uint32_t { return uniform(W.r.y); };
    auto isBig    = [&](const Window& W) -> bool { return firstEnd(W) - W.P > uint32_t(kWave); };
    // lanes of whole buckets: a prefix of the lanes (lane 0's bucket fits unless the window is big)
    auto wholeLane = [&](const Window& W) -> bool { return W.r.y != 0u && W.r.y <= W.P + uint32_t(kWave); };
    // the start of the window after W (end when W is the last)
    auto after = [&](const Window& W) -> uint32_t {
        if (W.P >= end) return end;
        if (isBig(W)) return firstEnd(W);
        const int lastWhole = __popcll(ballot(wholeLane(W))) - 1;
        return uint32_t(__builtin_amdgcn_readlane(int(W.r.y), lastWhole));
    };
    struct Props
    {
        T x, y, z, h;
    };
    auto gather = [&](const Window& W) -> Props {
        Props v{T(0), T(0), T(0), T(0)};
        if (W.P < end && !isBig(W) && wholeLane(W))
        {
            v.x = x[W.id];
            v.y = y[W.id];
            v.z = z[W.id];
            v.h = h[W.id];
        }
        return v;
    };

    Window W0{}, W1{}, W2{};
    if (off[0] >= end) return;
    open(W0, off[0]);
    const uint32_t p1 = after(W0);
    if (p1 < end) open(W1, p1);
    else W1.P = end;
    Props v0 = gather(W0);
    while (W0.P < end)
    {
        const uint32_t p2 = after(W1);
        if (p2 < end) open(W2, p2);
        else W2.P = end;
        const Props v1 = gather(W1);
        const uint32_t P = W0.P;
        if (isBig(W0))
        {
            // a bucket of more than 64 elements, [P, firstEnd)
            const uint32_t cnt = firstEnd(W0) - P;
            for (uint32_t c0 = 0; c0 < cnt; c0 += kWave)
            {
                const bool mine    = c0 + uint32_t(lane) < cnt;
                const K kc         = mine ? keysIn[P + c0 + lane] : K(0);
                const uint32_t idc = mine ? idxIn[P + c0 + lane] : 0u;
                uint32_t rank      = 0;
                for (uint32_t c1 = 0; c1 < cnt; c1 += kWave)
                {
                    const uint32_t m = c1 + uint32_t(lane);
                    waveSync(); // the previous chunk's reads are done
                    if (m < cnt)
                    {
                        wKey[lane] = keysIn[P + m];
                        wIdx[lane] = idxIn[P + m];
                    }
                    waveSync();
                    const uint32_t lim = cnt - c1 < uint32_t(kWave) ? cnt - c1 : uint32_t(kWave);
                    for (uint32_t q = 0; q < lim; ++q)
                    {
                        const K kq        = wKey[q];
                        const uint32_t iq = wIdx[q];
                        rank += uint32_t(kq < kc || (kq == kc && iq < idc));
                    }
                }
                if (mine) put(P + rank, kc, idc);
            }
        }
        else
        {
            const bool whole = wholeLane(W0);
            waveSync();
            wKey[lane] = W0.k;
            wIdx[lane] = W0.id;
            waveSync();
            // the whole buckets' keys ascend from bucket to bucket, so when every whole lane follows its left
            // neighbor in (key, index) order, lane order is the sorted order and the ranking is skipped: the common
            // case, since the keys arrive in the previous step's order and a bucket's slots follow lane order within
            // each wave
            const int left         = lane > 0 ? lane - 1 : 0;
            const uint32_t follows = lane > 0 ? addIfPairLess(0u, wKey[left], wIdx[left], W0.k, W0.id) : 1u;
            const bool inOrder     = ballot(whole && follows == 0u) == 0;
            if (whole)
            {
                uint32_t pos = P + uint32_t(lane);
                if (!inOrder)
                {
                    uint32_t rank = 0;
                    for (uint32_t q = W0.r.x - P; q < W0.r.y - P; ++q)
                    {
                        const K kq        = wKey[q];
                        const uint32_t iq = wIdx[q];
                        rank              = addIfPairLess(rank, kq, iq, W0.k, W0.id);
                    }
                    pos = W0.r.x + rank;
                }
                keysOut[pos] = W0.k;
                permOut[pos] = W0.id;
                xo[pos]      = v0.x;
                yo[pos]      = v0.y;
                zo[pos]      = v0.z;
                ho[pos]      = v0.h;
            }
        }
        W0 = W1;
        W1 = W2;
        v0 = v1;
    }
}

// ranking a bucket of s elements costs s * s / 64 compares in one wave; a larger bucket than this leaves the sort to
// the radix sort
constexpr uint32_t kMaxBucket = 2048;

// keys[0..n) are sorted in place (stable, i.e. by (key, input index)), perm[i] is the input index of sorted element i
// and xo..ho[i] = x..h[perm[i]]; leaves[0..L] is a cornerstone leaf array (leaves[0] = 0, L > 0). Keys at or above
// leaves[L - 1] go to the last bucket. keyBuf and bucketBuf hold n elements; segWs and scanWs grow as needed. If some
// bucket holds more than kMaxBucket keys, nothing is written but *done = false (one stream synchronization reads the
// largest bucket). prevLayout (optional): the leaf layout of these leaves over the previous positions of exactly n
// particles (keys[i] being the moved particle that sat at position i): it gives each wave of keys a leaf to start its
// bucket search from (waveLeafHintKernel, into hintWs). maxOut (optional, host): receives the largest bucket. With
// deferMax (maxOut then in pinned memory) it is copied there asynchronously and the bucket path is taken without
// reading it first (no synchronization): the result is the same for any bucket size (large buckets are ranked in
// chunks), the guard only bounds the work, so a caller that expects small buckets (they were small in its previous
// step) reads *maxOut at its next synchronization and guards the step after that.
template<class K, class T>
hipcstone_status sortByLeafBuckets(hipcstone_stream s, K* keys, size_t n, const K* leaves, int32_t L, const T* x,
                                   const T* y, const T* z, const T* h, T* xo, T* yo, T* zo, T* ho, uint32_t* perm,
                                   K* keyBuf, uint32_t* bucketBuf, DeviceBuf& segWs, DeviceBuf& scanWs, bool* done,
                                   const uint32_t* prevLayout = nullptr, DeviceBuf* hintWs = nullptr,
                                   uint32_t* maxOut = nullptr, bool deferMax = false)
{
    *done = false;
    if (n == 0)
    {
        *done = true;
        return HIPCSTONE_OK;
    }
    if (L <= 0) return invalid("sortByLeafBuckets: empty leaf array");
    if (auto e = segWs.reserve((2 * n + 2 * size_t(L) + 4) * sizeof(uint32_t))) return e;
    if (auto e = scanWs.reserve(hipcstone_scan_workspace_size(size_t(L)))) return e;
    uint32_t* segCount  = segWs.as<uint32_t>();
    uint32_t* segOffset = segCount + L; // L + 1 entries
    uint32_t* slot      = segOffset + L + 1;
    uint32_t* idx       = slot + n;

    hipStream_t st     = stream(s);
    const uint32_t n32 = uint32_t(n);

    uint32_t* maxBucket = idx + n;

    zeroU32Kernel<<<ceilDiv(size_t(L), 256), 256, 0, st>>>(segCount, uint32_t(L));
    HC_CHECK_LAUNCH("zeroU32Kernel");
    zeroU32Kernel<<<1, 256, 0, st>>>(maxBucket, 1u);
    HC_CHECK_LAUNCH("zeroU32Kernel");
    // the previous layout (of the same particle count, in the order the keys come in) gives each wave a leaf to
    // start from
    uint32_t* hint = nullptr;
    if (prevLayout && hintWs)
    {
        if (auto e = hintWs->reserve(size_t(ceilDiv(n, 64)) * sizeof(uint32_t))) return e;
        hint = hintWs->as<uint32_t>();
        waveLeafHintKernel<<<ceilDiv(size_t(L), 256), 256, 0, st>>>(prevLayout, L, n32, hint);
        HC_CHECK_LAUNCH("waveLeafHintKernel");
    }
    bucketByLeafKernel<K><<<ceilDiv(n, 256), 256, 0, st>>>(keys, n32, leaves, L, segCount, bucketBuf, slot, hint);
    HC_CHECK_LAUNCH("bucketByLeafKernel");
    maxBucketKernel<<<ceilDiv(size_t(L), kMaxPerBlock), 256, 0, st>>>(segCount, uint32_t(L), maxBucket);
    HC_CHECK_LAUNCH("maxBucketKernel");
    if (maxOut && deferMax) { HC_CHECK(hipMemcpyAsync(maxOut, maxBucket, sizeof(uint32_t), hipMemcpyDeviceToHost, st)); }
    else
    {
        uint32_t largest = 0;
        HC_CHECK(hipMemcpyAsync(&largest, maxBucket, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HC_CHECK(hipStreamSynchronize(st));
        if (maxOut) *maxOut = largest;
        if (largest > kMaxBucket) return HIPCSTONE_OK;
    }
    if (auto e = hipcstone_leaf_layout(s, segCount, L, segOffset, scanWs.ptr, scanWs.bytes)) return e;
    scatterByBucketKernel<K><<<ceilDiv(n, 256), 256, 0, st>>>(keys, n32, bucketBuf, slot, segOffset, keyBuf, idx);
    HC_CHECK_LAUNCH("scatterByBucketKernel");
    const unsigned blocks = ceilDiv(ceilDiv(size_t(L), kBucketsPerWave), 4);
    sortBucketsGatherKernel<K, T><<<blocks, 256, 0, st>>>(keyBuf, idx, segOffset, L, keys, perm, x, y, z, h, xo, yo,
                                                          zo, ho);
    HC_CHECK_LAUNCH("sortBucketsGatherKernel");
    *done = true;
    return HIPCSTONE_OK;
}

} // namespace hipcstone
