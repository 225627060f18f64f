// a18: one-rank Domain::sync on the GPU.
//
// Reference semantics (paths under include/cstone/): domain/domain.hpp:169-218 (sync), domain/assignment.hpp:54-123
// (global tree: fitted / shrink-limited box, keys, stable sort, the update rule), sfc/box_mpi.hpp:51-105
// (makeGlobalBox) and sfc/box.hpp:398-415 (limitBoxShrinking), focus/octree_focus_mpi.hpp:61-87, :101-176,
// :193-245, :585-603 (focus tree: construction, updateTree, updateCounts, converge), focus/octree_focus.hpp:64-111
// (CombinedUpdate::updateFocus), focus/rebalance.hpp:31-252 (mergeCountAndMacOp, protectAncestors, enforceKeys),
// focus/inject.hpp:21-45 (injectKeys).
//
// With one rank the focus area is the whole key range and every global leaf key is mandatory in the focus tree.
// Then MAC flags never enter a decision (mergeCountAndMacOp: every sibling group lies in the focus "fringe" and
// every leaf in the focus), halo discovery finds nothing outside the assigned range, and the particle layout is the
// scan of the focus leaf counts. Keys are Hilbert keys, as the reference's Domain computes them (SfcKind).
#include <algorithm>
#include <memory>
#include <vector>

#include "common.hpp"
#include "host_util.hpp"
#include "runtime.hpp"
#include "scan.hpp"
#include "bucket_sort.hpp"

namespace hipcstone
{

// ---------------------------------------------------------------- coordinate extrema (makeGlobalBox's minMax)

constexpr int kMmThreads = 256;
constexpr int kMmBlocks  = 1024;

template<class T>
__global__ __launch_bounds__(kMmThreads) void minMaxPartialKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                                  const T* __restrict__ z, size_t n,
                                                                  T* __restrict__ partials)
{
    const T* a[3] = {x, y, z};
    __shared__ T sh[6][kMmThreads / kWave];
    const int lane = laneId(), w = threadIdx.x / kWave;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T lo = a[d][0], hi = a[d][0];
        for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        {
            T v = a[d][i];
            lo  = v < lo ? v : lo;
            hi  = v > hi ? v : hi;
        }
        for (int off = kWave / 2; off > 0; off >>= 1)
        {
            T o = __shfl_xor(lo, off, kWave);
            lo  = o < lo ? o : lo;
            o   = __shfl_xor(hi, off, kWave);
            hi  = o > hi ? o : hi;
        }
        if (lane == 0)
        {
            sh[2 * d][w]     = lo;
            sh[2 * d + 1][w] = hi;
        }
    }
    __syncthreads();
    if (threadIdx.x < 6)
    {
        T v = sh[threadIdx.x][0];
        for (int k = 1; k < kMmThreads / kWave; ++k)
        {
            T o = sh[threadIdx.x][k];
            v   = (threadIdx.x & 1) ? (o > v ? o : v) : (o < v ? o : v);
        }
        partials[size_t(blockIdx.x) * 6 + threadIdx.x] = v;
    }
}

template<class T>
__global__ __launch_bounds__(6 * kWave) void minMaxFinalKernel(const T* __restrict__ partials, int numPartials,
                                                                T* __restrict__ out)
{
    // wave c reduces component c (xmin, xmax, ymin, ...) of the per-block partials: strided per lane, then across
    // the wave, as the partials themselves were reduced
    const int c = threadIdx.x / kWave, lane = laneId();
    const bool isMax = c & 1;
    T v = partials[c];
    for (int b = lane; b < numPartials; b += kWave)
    {
        T o = partials[size_t(b) * 6 + c];
        v   = isMax ? (o > v ? o : v) : (o < v ? o : v);
    }
    for (int off = kWave / 2; off > 0; off >>= 1)
    {
        T o = __shfl_xor(v, off, kWave);
        v   = isMax ? (o > v ? o : v) : (o < v ? o : v);
    }
    if (lane == 0) out[c] = v;
}

// ---------------------------------------------------------------- global tree: largest leaf count

__global__ void maxCountKernel(const uint32_t* __restrict__ counts, int32_t n, uint32_t* __restrict__ out)
{
    uint32_t m = 0;
    for (int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x); i < n; i += int32_t(gridDim.x * blockDim.x))
        m = counts[i] > m ? counts[i] : m;
    for (int off = kWave / 2; off > 0; off >>= 1)
    {
        uint32_t o = __shfl_xor(m, off, kWave);
        m          = o > m ? o : m;
    }
    // the maximum only grows: skip the atomic when it already holds at least m (thousands of waves share *out)
    if (laneId() == 0 && m > __hip_atomic_load(out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(out, m);
}

// ---------------------------------------------------------------- focus tree decisions

template<class K>
__device__ inline int lastNzPlace(K x)
{
    if (x == 0) return int(maxLevel<K>());
    int tz = sizeof(K) == 8 ? __builtin_ctzll(uint64_t(x)) : __builtin_ctz(uint32_t(x));
    return int(maxLevel<K>()) - tz / 3;
}

template<class K>
__device__ inline K makePrefix(K a)
{
    if (a == 0) return K(1);
    return encodePlaceholderBit(a, 3 * lastNzPlace(a));
}

// tree/octree.hpp:199-215 containingNode
template<class K>
__device__ inline int32_t containingNode(K nodeKey, const K* prefixes, const int32_t* childOffsets)
{
    int nodeLevel = int(decodePrefixLength(nodeKey)) / 3;
    K key         = decodePlaceholderBit(nodeKey);
    int32_t ret   = 0;
    for (int i = 1; i <= nodeLevel; ++i)
    {
        if (childOffsets[ret] == 0 || nodeKey == prefixes[ret]) break;
        ret = childOffsets[ret] + int32_t(octalDigit(key, unsigned(i)));
    }
    return ret;
}

// focus/rebalance.hpp:31-73 mergeCountAndMacOp for every node; macs may be NULL (all 0)
template<class K>
__global__ void focusNodeOpsKernel(const K* __restrict__ prefixes, const int32_t* __restrict__ childOffsets,
                                   const int32_t* __restrict__ parents, const uint32_t* __restrict__ counts,
                                   const uint8_t* __restrict__ macs, int32_t numNodes, K focusStart, K focusEnd,
                                   uint32_t bucketSize, int32_t* __restrict__ ops)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= numNodes) return;
    K nodeKey      = prefixes[i];
    unsigned level = decodePrefixLength(nodeKey) / 3;
    if (i > 0)
    {
        int32_t parent    = parents[(i - 1) / 8];
        bool countMerge   = counts[parent] <= bucketSize;
        bool macMerge     = !macs || macs[parent] == 0;
        K firstGroupKey   = decodePlaceholderBit(prefixes[parent]);
        K lastGroupKey    = firstGroupKey + K(8) * nodeRange<K>(level);
        bool inFringe     = lastGroupKey > focusStart && focusEnd > firstGroupKey;
        if (countMerge || (macMerge && !inFringe))
        {
            ops[i] = 0;
            return;
        }
    }
    K nodeStart  = decodePlaceholderBit(nodeKey);
    bool isLeaf  = childOffsets[i] == 0;
    bool inFocus = nodeStart >= focusStart && nodeStart < focusEnd;
    int op       = 1;
    if (isLeaf && ((macs && macs[i]) || inFocus))
    {
        uint32_t c = counts[i];
        if (level + 3 < maxLevel<K>() && c > 4096u * bucketSize) op = 4096;
        else if (level + 2 < maxLevel<K>() && c > 512u * bucketSize) op = 512;
        else if (level + 1 < maxLevel<K>() && c > 64u * bucketSize) op = 64;
        else if (level < maxLevel<K>() && c > bucketSize) op = 8;
    }
    ops[i] = op;
}

// focus/rebalance.hpp:186-252 enforceKeys, one key per lane. Clearing merges only turns 0 into 1 and splitting
// takes a maximum, so the final ops do not depend on the order of the keys; a key that sees a merge already
// cancelled by another key finds its whole ancestor chain cancelled by that key, and that key reports at least
// cancelMerge, so the maximum status does not depend on the order either.
template<class K>
__global__ void enforceKeysKernel(const K* __restrict__ keys, int32_t numKeys, const K* __restrict__ prefixes,
                                  const int32_t* __restrict__ childOffsets, const int32_t* __restrict__ parents,
                                  int32_t* ops, int32_t* status)
{
    int32_t k = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= numKeys) return;
    K key = keys[k];
    if (key == 0 || key == nodeRange<K>(0)) return;
    K want             = makePrefix(key);
    int32_t nodeIdx    = containingNode(want, prefixes, childOffsets);
    K have             = prefixes[nodeIdx];
    unsigned levelHave = decodePrefixLength(have) / 3;
    bool trySplit      = have != want && levelHave < maxLevel<K>();
    int cur            = __hip_atomic_load(ops + nodeIdx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool undoMerges    = cur == 0 || trySplit;
    int st             = 0;
    if (undoMerges && nodeIdx > 0)
    {
        st             = 1;
        int32_t parent = nodeIdx;
        for (int guard = 0; guard <= int(maxLevel<K>()) + 1; ++guard)
        {
            parent     = parents[(parent - 1) / 8];
            int32_t fs = childOffsets[parent];
            for (int32_t i = fs; i < fs + 8; ++i)
                atomicCAS(ops + i, 0, 1);
            if (parent == 0) break;
        }
    }
    // ops and status only grow during this kernel (compare-and-swap 0 -> 1, atomic max), so an atomic max whose
    // value the target already holds is skipped: on a root-only focus tree every key would otherwise hit ops[0]
    if (trySplit)
    {
        int levelDiff = lastNzPlace(key) - int(levelHave);
        st            = levelDiff > 1 ? 3 : 2;
        levelDiff     = levelDiff < 1 ? levelDiff : 1;
        const int v   = 1 << (3 * levelDiff);
        if (__hip_atomic_load(ops + nodeIdx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v)
            atomicMax(ops + nodeIdx, v);
    }
    if (st && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < st) atomicMax(status, st);
}

// focus/rebalance.hpp:90-107 nzAncestorOp into a separate array (the reference's in-place parallel update reads
// the same values: a rewritten ancestor only ever takes the op of its own closest non-zero ancestor when it is that
// ancestor's left-most descendant, which leaves every descendant's answer unchanged); changes is set to 1 if
// some op differs from 1
template<class K>
__global__ void protectAncestorsKernel(const K* __restrict__ prefixes, const int32_t* __restrict__ parents,
                                       const int32_t* __restrict__ in, int32_t numNodes, int32_t* __restrict__ out,
                                       int32_t* changes)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    bool notOne = false;
    if (i < numNodes)
    {
        int32_t op;
        if (i == 0) op = in[0];
        else
        {
            int32_t a = i;
            for (int guard = 0; in[a] == 0 && guard <= int(maxLevel<K>()) + 1; ++guard)
                a = parents[(a - 1) / 8];
            op = decodePlaceholderBit(prefixes[i]) == decodePlaceholderBit(prefixes[a]) ? in[a] : 0;
        }
        out[i] = op;
        notOne = op != 1;
    }
    // only whether some op differs from 1 is used: set the flag once instead of counting (one atomic per wave on a
    // single address serializes over millions of nodes)
    const uint64_t b = ballot(notOne);
    if (laneId() == 0 && b && __hip_atomic_load(changes, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __hip_atomic_store(changes, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void gatherLeafOpsKernel(const int32_t* __restrict__ nodeOps, const int32_t* __restrict__ leafToInternal,
                                    int32_t numLeaves, int32_t* __restrict__ leafOps, int32_t* anyNotOne)
{
    int32_t l   = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    bool notOne = false;
    if (l < numLeaves)
    {
        int32_t op = nodeOps[leafToInternal[l]];
        leafOps[l] = op;
        notOne     = op != 1;
    }
    else if (l == numLeaves) { leafOps[l] = 0; }
    const uint64_t b = ballot(notOne);
    if (laneId() == 0 && b && __hip_atomic_load(anyNotOne, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        __hip_atomic_store(anyNotOne, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// node counts of a linked tree from the sorted particle keys: the keys in [start, start + range) of each node, as
// computeNodeCounts gives for leaves and the count upsweep (NodeCount) gives for internal nodes
template<class K>
__device__ inline size_t lowerBoundKeys(const K* keys, size_t n, K v)
{
    size_t lo = 0, len = n;
    while (len > 0)
    {
        size_t half = len >> 1;
        if (keys[lo + half] < v)
        {
            lo += half + 1;
            len -= half + 1;
        }
        else { len = half; }
    }
    return lo;
}

template<class K>
__global__ void nodeKeyCountsKernel(const K* __restrict__ prefixes, int32_t numNodes, const K* __restrict__ keys,
                                    size_t numKeys, uint32_t* __restrict__ counts)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= numNodes) return;
    K pfx       = prefixes[i];
    unsigned pl = decodePrefixLength(pfx);
    K start     = decodePlaceholderBit(pfx);
    K end       = start + nodeRange<K>(pl / 3);
    size_t c    = lowerBoundKeys(keys, numKeys, end) - lowerBoundKeys(keys, numKeys, start);
    counts[i]   = c < 0xFFFFFFFFull ? uint32_t(c) : 0xFFFFFFFFu;
}

__global__ void gatherU32Kernel(const uint32_t* __restrict__ src, const int32_t* __restrict__ idx, int32_t n,
                                uint32_t* __restrict__ dst)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) dst[i] = src[idx[i]];
}

// injectKeys: keys sorted; flag the first of each run of equal keys
template<class K>
__global__ void uniqueFlagsKernel(const K* __restrict__ keys, int32_t n, uint32_t* __restrict__ flags)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

template<class K>
__global__ void uniqueScatterKernel(const K* __restrict__ keys, const uint32_t* __restrict__ pos, int32_t n,
                                    K* __restrict__ out)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n && (i == 0 || keys[i] != keys[i - 1])) out[pos[i]] = keys[i];
}

template<class K>
__global__ void setKeysKernel(K* __restrict__ dst, K a, K b)
{
    dst[0] = a;
    dst[1] = b;
}

__global__ void setU32Kernel(uint32_t* __restrict__ dst, uint32_t v) { dst[0] = v; }

__global__ void zeroI32Kernel(int32_t* __restrict__ dst, int n)
{
    if (int(threadIdx.x) < n) dst[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- property gather (the sync's SFC order)

template<int E>
struct PropRecord
{
    uint8_t b[E];
};
template<>
struct alignas(4) PropRecord<4>
{
    uint32_t v;
};
template<>
struct alignas(8) PropRecord<8>
{
    uint64_t v;
};
template<>
struct alignas(16) PropRecord<16>
{
    uint64_t v[2];
};

template<int E>
__global__ void gatherPropKernel(const uint32_t* __restrict__ perm, size_t n, const PropRecord<E>* __restrict__ src,
                                 PropRecord<E>* __restrict__ dst)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// ---------------------------------------------------------------- the domain object

template<class K, class T>
struct DomainApi;

#define HC_DAPI(KS, K, RS, T)                                                                                          \
    template<>                                                                                                         \
    struct DomainApi<K, T>                                                                                             \
    {                                                                                                                  \
        using Box = hipcstone_box_##RS;                                                                                \
        static hipcstone_status keys(hipcstone_stream s, const T* x, const T* y, const T* z, K* k, size_t n,          \
                                     const Box* b)                                                                     \
        {                                                                                                              \
            return hipcstone_compute_sfc_keys_hilbert_##KS##_##RS(s, x, y, z, k, n, b, 0);                            \
        }                                                                                                              \
        static size_t sortWs(size_t n) { return hipcstone_sort_gather_workspace_size_##KS(n); }                        \
        static hipcstone_status sortGather(hipcstone_stream s, K* k, uint32_t* p, size_t n, const T* xi, const T* yi, \
                                           const T* zi, const T* hi, T* xo, T* yo, T* zo, T* ho, K* kb, uint32_t* vb, \
                                           void* ws, size_t wsb)                                                       \
        {                                                                                                              \
            return hipcstone_sort_gather_##KS##_##RS(s, k, p, n, xi, yi, zi, hi, xo, yo, zo, ho, kb, vb, ws, wsb);    \
        }                                                                                                              \
        static size_t keySortWs(size_t n) { return hipcstone_sort_by_key_workspace_size_##KS(n); }                     \
        static hipcstone_status sortKeys(hipcstone_stream s, K* k, size_t n, K* kb, void* ws, size_t wsb)             \
        {                                                                                                              \
            return hipcstone_sort_by_key_##KS(s, k, nullptr, n, kb, nullptr, ws, wsb, 0);                              \
        }                                                                                                              \
        static hipcstone_status update(hipcstone_stream s, const K* keys, size_t n, uint32_t bucket, const K* tree,   \
                                       const uint32_t* counts, int32_t* nn, K* nt, uint32_t* nc, int32_t cap,          \
                                       void* ws, size_t wsb, int* conv)                                                \
        {                                                                                                              \
            return hipcstone_update_octree_##KS(s, keys, n, bucket, tree, counts, nn, nt, nc, cap, 0xFFFFFFFFu, ws,   \
                                                wsb, conv);                                                            \
        }                                                                                                              \
        static hipcstone_status rebalance(hipcstone_stream s, const K* tree, int32_t nn, int32_t newNn,               \
                                          const int32_t* ops, K* newTree)                                              \
        {                                                                                                              \
            return hipcstone_rebalance_tree_##KS(s, tree, nn, newNn, ops, newTree);                                    \
        }                                                                                                              \
        static size_t buildWs(int32_t L) { return hipcstone_build_octree_workspace_size_##KS(L); }                    \
        static hipcstone_status build(hipcstone_stream s, const K* leaves, int32_t L, K* pf, int32_t* co,             \
                                      int32_t* par, int32_t* lr, int32_t* i2l, int32_t* l2i, void* ws, size_t wsb,     \
                                      int32_t* lrh)                                                                    \
        {                                                                                                              \
            return hipcstone_build_octree_##KS(s, leaves, L, pf, co, par, lr, i2l, l2i, ws, wsb, lrh);                \
        }                                                                                                              \
        static hipcstone_status centers(hipcstone_stream s, const K* pf, int32_t N, T* c, T* sz, const Box* b)        \
        {                                                                                                              \
            return hipcstone_compute_geo_centers_##KS##_##RS(s, pf, N, c, sz, b);                                     \
        }                                                                                                              \
        static hipcstone_status leafCenters(hipcstone_stream s, const T* x, const T* y, const T* z, const T* m,       \
                                            const int32_t* l2i, int32_t L, const uint32_t* layout, T* c)               \
        {                                                                                                              \
            return hipcstone_leaf_mass_centers_##RS(s, x, y, z, m, l2i, L, layout, c);                                \
        }                                                                                                              \
        static hipcstone_status upsweep(hipcstone_stream s, int32_t levels, const int32_t* lrh, const int32_t* co,    \
                                        T* c)                                                                          \
        {                                                                                                              \
            return hipcstone_upsweep_centers_##RS(s, levels, lrh, co, c);                                             \
        }                                                                                                              \
        static hipcstone_status setMac(hipcstone_stream s, const K* pf, int32_t N, T* c, float invTheta,             \
                                       const Box* b)                                                                   \
        {                                                                                                              \
            return hipcstone_set_mac_##KS##_##RS(s, pf, N, c, invTheta, b);                                           \
        }                                                                                                              \
        static void makeBox(const T* lo, const T* hi, const uint8_t* bnd, Box* b)                                      \
        {                                                                                                              \
            hipcstone_make_box_##RS(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], bnd[0], bnd[1], bnd[2], b);             \
        }                                                                                                              \
    };

HC_DAPI(u32, uint32_t, f32, float)
HC_DAPI(u32, uint32_t, f64, double)
HC_DAPI(u64, uint64_t, f32, float)
HC_DAPI(u64, uint64_t, f64, double)
#undef HC_DAPI

struct DomainBase
{
    StageTimer<HIPCSTONE_DOMAIN_STAGES> timer;
    virtual ~DomainBase() = default;
    virtual hipcstone_status sync(hipcstone_stream s, void* x, void* y, void* z, void* h, size_t n,
                                  void* keysOut) = 0;
    virtual void view(hipcstone_domain_view* v) const = 0;
    virtual hipcstone_status gatherProperties(hipcstone_stream s, void* const* arrays, const size_t* elemBytes,
                                              int numArrays) = 0;
    virtual hipcstone_status syncGrav(hipcstone_stream s, void* x, void* y, void* z, void* h, void* m, size_t n,
                                      void* keysOut) = 0;
    virtual hipcstone_status updateExpansionCenters(hipcstone_stream s, const void* x, const void* y, const void* z,
                                                    const void* m) = 0;
};

template<class K, class T>
class Domain final : public DomainBase
{
    using A   = DomainApi<K, T>;
    using Box = typename A::Box;

public:
    Domain(uint32_t bucket, uint32_t bucketFocus, float theta, const Box& box)
        : bucket_(bucket)
        , bucketFocus_(bucketFocus)
        , invTheta_(float(1.0 / double(theta))) // the reference passes 1.0 / theta_ (a double) to a float argument
    {
        for (int d = 0; d < 3; ++d)
        {
            lo_[d]  = box.lim[2 * d];
            hi_[d]  = box.lim[2 * d + 1];
            bnd_[d] = box.bnd[d];
        }
        A::makeBox(lo_, hi_, bnd_, &box_);
    }

    hipcstone_status sync(hipcstone_stream s, void* xv, void* yv, void* zv, void* hv, size_t n,
                          void* keysOut) override
    {
        if (n == 0) return invalid("domain sync: no particles");
        if (n > (size_t(1) << 30) - 1) return invalid("domain sync: more than 2^30-1 particles");
        if (!xv || !yv || !zv || !hv) return invalid("domain sync: null particle array");
        T* x = static_cast<T*>(xv);
        T* y = static_cast<T*>(yv);
        T* z = static_cast<T*>(zv);
        T* h = static_cast<T*>(hv);
        if (auto e = reserveParticles(n)) return e;
        const size_t prevN = n_;
        n_                 = n;
        auto& tm = this->timer;
#define HC_STAGE(id, call)                                                                                             \
    do {                                                                                                               \
        if (auto e_ = tm.begin(id, s)) return e_;                                                                     \
        if (auto e_ = (call)) return e_;                                                                               \
        if (auto e_ = tm.end(s)) return e_;                                                                            \
    } while (0)
        HC_STAGE(0, updateBox(s, x, y, z));
        HC_STAGE(1, A::keys(s, x, y, z, keys_.as<K>(), n, &box_));
        HC_STAGE(2, sortAndGather(s, x, y, z, h, n, prevN));
        HC_STAGE(3, globalUpdate(s));
        if (first_)
        {
            if (auto e = initFocus(s)) return e;
            int conv = 0;
            focusSteps_ = 0;
            while (!conv)
            {
                HC_STAGE(4, focusUpdate(s, &conv));
                HC_STAGE(5, focusCounts(s));
                if (++focusSteps_ > 64) return invalid("domain sync: focus tree did not converge in 64 updates");
            }
        }
        else { focusSteps_ = 0; }
        int conv = 0;
        HC_STAGE(4, focusUpdate(s, &conv));
        HC_STAGE(5, focusCounts(s));
        focusSteps_++;
        focusConverged_ = conv;
        HC_STAGE(6, finish(s, x, y, z, h, keysOut));
#undef HC_STAGE
        first_          = false;
        centersCurrent_ = false;
        return HIPCSTONE_OK;
    }

    // Domain::syncGrav at one rank (domain.hpp:220-287): the MAC flags decide nothing (no node lies outside the
    // focus), so the particle order and both trees are those of sync; m is reordered with x, y, z, h, and the focus
    // tree's expansion centers are the mass centers of its final tree (updateCenters, octree_focus_mpi.hpp:351-414)
    // with the vector-MAC radius of 1 / theta (updateMacs -> setMacRadius, :446-453)
    hipcstone_status syncGrav(hipcstone_stream s, void* x, void* y, void* z, void* h, void* m, size_t n,
                              void* keysOut) override
    {
        if (!m) return invalid("domain syncGrav: null mass array");
        if (auto e = sync(s, x, y, z, h, n, keysOut)) return e;
        void* arrays[1]        = {m};
        const size_t bytes[1] = {sizeof(T)};
        if (auto e = gatherProperties(s, arrays, bytes, 1)) return e;
        return updateExpansionCenters(s, x, y, z, m);
    }

    // Domain::updateExpansionCenters (domain.hpp:366-372): leaf mass centers over the focus layout, the upsweep, and
    // setMacRadius(1 / theta); x, y, z, m are in the order of the last sync
    hipcstone_status updateExpansionCenters(hipcstone_stream s, const void* xv, const void* yv, const void* zv,
                                            const void* mv) override
    {
        if (n_ == 0) return invalid("domain expansion centers: no sync yet");
        if (!xv || !yv || !zv || !mv) return invalid("domain expansion centers: null array");
        if (auto e = expCenters_.reserve(size_t(fN_) * 4 * sizeof(T))) return e;
        if (auto e = this->timer.begin(7, s)) return e;
        const int32_t NI = fN_ - fL_;
        if (auto e = A::leafCenters(s, static_cast<const T*>(xv), static_cast<const T*>(yv),
                                    static_cast<const T*>(zv), static_cast<const T*>(mv),
                                    leafToInternal_.as<int32_t>() + NI, fL_, layout_.as<uint32_t>(),
                                    expCenters_.as<T>()))
            return e;
        int32_t lr[32];
        constexpr int32_t levels = int32_t(maxLevel<K>());
        HC_CHECK(hipMemcpyAsync(lr, levelRange_.ptr, (levels + 2) * sizeof(int32_t), hipMemcpyDeviceToHost,
                                stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s)));
        if (auto e = A::upsweep(s, levels, lr, childOffsets_.as<int32_t>(), expCenters_.as<T>())) return e;
        if (auto e = A::setMac(s, prefixes_.as<K>(), fN_, expCenters_.as<T>(), invTheta_, &box_)) return e;
        if (auto e = this->timer.end(s)) return e;
        centersCurrent_ = true;
        return HIPCSTONE_OK;
    }

    void view(hipcstone_domain_view* v) const override
    {
        v->num_particles     = n_;
        v->start_index       = 0;
        v->end_index         = n_;
        v->num_global_leaves = numGlobal_;
        v->num_leaves        = fL_;
        v->num_nodes         = fN_;
        v->num_internal      = fN_ - fL_;
        v->global_steps      = globalSteps_;
        v->focus_steps       = focusSteps_;
        v->focus_converged   = focusConverged_;
        for (int d = 0; d < 3; ++d)
        {
            v->box_lo[d] = double(lo_[d]);
            v->box_hi[d] = double(hi_[d]);
        }
        v->keys             = keys_.ptr;
        v->global_leaves    = gTree_[gCur_].ptr;
        v->global_counts    = gCounts_[gCur_].ptr;
        v->leaves           = fLeaves_.ptr;
        v->counts           = fLeafCounts_.ptr;
        v->node_counts      = fNodeCounts_.ptr;
        v->prefixes         = prefixes_.ptr;
        v->child_offsets    = childOffsets_.ptr;
        v->parents          = parents_.ptr;
        v->internal_to_leaf = internalToLeaf_.ptr;
        v->leaf_to_internal = leafToInternal_.ptr;
        v->level_range      = levelRange_.ptr;
        v->centers          = centers_.ptr;
        v->sizes            = sizes_.ptr;
        v->layout           = layout_.ptr;
        v->expansion_centers = centersCurrent_ ? expCenters_.ptr : nullptr;
    }

    // reorder property arrays (n elements each, in the order the last sync received x, y, z, h) into the SFC order
    // of that sync, in place through one scratch buffer
    hipcstone_status gatherProperties(hipcstone_stream s, void* const* arrays, const size_t* elemBytes,
                                      int numArrays) override
    {
        if (n_ == 0) return invalid("domain gather: no sync yet");
        if (numArrays < 0 || (numArrays > 0 && (!arrays || !elemBytes))) return invalid("domain gather: null list");
        for (int a = 0; a < numArrays; ++a)
        {
            size_t e = elemBytes[a];
            if (!arrays[a]) return invalid("domain gather: null array");
            if (e != 1 && e != 2 && e != 4 && e != 8 && e != 12 && e != 16)
                return invalid("domain gather: element size must be 1, 2, 4, 8, 12 or 16 bytes");
            if (auto r = propScratch_.reserve(n_ * e)) return r;
            unsigned blocks = ceilDiv(n_, 256);
            const uint32_t* perm = perm_.as<uint32_t>();
            switch (e)
            {
                case 1: gatherPropKernel<1><<<blocks, 256, 0, stream(s)>>>(perm, n_, static_cast<const PropRecord<1>*>(arrays[a]), propScratch_.as<PropRecord<1>>()); break;
                case 2: gatherPropKernel<2><<<blocks, 256, 0, stream(s)>>>(perm, n_, static_cast<const PropRecord<2>*>(arrays[a]), propScratch_.as<PropRecord<2>>()); break;
                case 4: gatherPropKernel<4><<<blocks, 256, 0, stream(s)>>>(perm, n_, static_cast<const PropRecord<4>*>(arrays[a]), propScratch_.as<PropRecord<4>>()); break;
                case 8: gatherPropKernel<8><<<blocks, 256, 0, stream(s)>>>(perm, n_, static_cast<const PropRecord<8>*>(arrays[a]), propScratch_.as<PropRecord<8>>()); break;
                case 12: gatherPropKernel<12><<<blocks, 256, 0, stream(s)>>>(perm, n_, static_cast<const PropRecord<12>*>(arrays[a]), propScratch_.as<PropRecord<12>>()); break;
                default: gatherPropKernel<16><<<blocks, 256, 0, stream(s)>>>(perm, n_, static_cast<const PropRecord<16>*>(arrays[a]), propScratch_.as<PropRecord<16>>()); break;
            }
            HC_CHECK_LAUNCH("gatherPropKernel");
            HC_CHECK(hipMemcpyAsync(arrays[a], propScratch_.ptr, n_ * e, hipMemcpyDeviceToDevice, stream(s)));
        }
        return HIPCSTONE_OK;
    }

private:
    // makeGlobalBox at one rank, then limitBoxShrinking after the first call
    hipcstone_status updateBox(hipcstone_stream s, const T* x, const T* y, const T* z)
    {
        bool keep[3];
        for (int d = 0; d < 3; ++d)
            keep[d] = bnd_[d] == HIPCSTONE_PERIODIC || bnd_[d] == HIPCSTONE_FIXED;
        T ext[6] = {lo_[0], hi_[0], lo_[1], hi_[1], lo_[2], hi_[2]};
        if (!keep[0] || !keep[1] || !keep[2])
        {
            int blocks = int(std::min<size_t>(kMmBlocks, ceilDiv(n_, kMmThreads)));
            minMaxPartialKernel<T><<<blocks, kMmThreads, 0, stream(s)>>>(x, y, z, n_, mmPartials_.as<T>());
            HC_CHECK_LAUNCH("minMaxPartialKernel");
            minMaxFinalKernel<T><<<1, 6 * kWave, 0, stream(s)>>>(mmPartials_.as<T>(), blocks, mmOut_.as<T>());
            HC_CHECK_LAUNCH("minMaxFinalKernel");
            T mm[6];
            HC_CHECK(hipMemcpyAsync(mm, mmOut_.ptr, sizeof(mm), hipMemcpyDeviceToHost, stream(s)));
            HC_CHECK(hipStreamSynchronize(stream(s)));
            for (int d = 0; d < 3; ++d)
                if (!keep[d])
                {
                    ext[2 * d]     = mm[2 * d];
                    ext[2 * d + 1] = mm[2 * d + 1];
                }
        }
        T side = std::max({T(ext[1] - ext[0]), T(ext[3] - ext[2]), T(ext[5] - ext[4])});
        for (int d = 0; d < 3; ++d)
            if (bnd_[d] == HIPCSTONE_CUBIC_OPEN) ext[2 * d + 1] = std::max(ext[2 * d + 1], T(ext[2 * d] + side));
        if (!first_)
        {
            const T shrink = T(0.05);
            for (int d = 0; d < 3; ++d)
            {
                T len          = box_.len[d];
                T limLo        = T(lo_[d] + T(shrink * len));
                T limHi        = T(hi_[d] - T(shrink * len));
                ext[2 * d]     = std::min(ext[2 * d], limLo);
                ext[2 * d + 1] = std::max(ext[2 * d + 1], limHi);
            }
        }
        for (int d = 0; d < 3; ++d)
        {
            lo_[d] = ext[2 * d];
            hi_[d] = ext[2 * d + 1];
        }
        A::makeBox(lo_, hi_, bnd_, &box_);
        return HIPCSTONE_OK;
    }

    // stable sort + gather of the call's particles into xyzh_ (one timed stage): later calls go through the buckets
    // of the previous global leaf array, which hands a step with a bucket above its limit to the radix sort (the
    // same result either way)
    // prevN: the particle count of the previous sync
    hipcstone_status sortAndGather(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, size_t n,
                                   size_t prevN)
    {
        bool bucketed = false;
        uint32_t* pin = pinnedStaging();
        if (!pin) return HIPCSTONE_HIP_ERROR;
        if (maxPending_)
        {
            lastMaxBucket_ = pin[kMaxWord]; // the previous sort's largest bucket, copied before its syncs
            maxPending_    = false;
        }
        if (numGlobal_ > 0)
        {
            // the global leaves' counts are exact (no cap), so with an unchanged particle count their layout is the
            // previous step's, over the previous positions of the particles the keys come in: a search hint per wave
            // (any input order sorts the same)
            const uint32_t* prevLayout = nullptr;
            if (n == prevN)
            {
                if (auto e = gLayout_.reserve((size_t(numGlobal_) + 1) * sizeof(uint32_t))) return e;
                if (auto e = bsScanWs_.reserve(hipcstone_scan_workspace_size(size_t(numGlobal_)))) return e;
                if (auto e = hipcstone_leaf_layout(s, gCounts_[gCur_].template as<uint32_t>(), numGlobal_,
                                                   gLayout_.as<uint32_t>(), bsScanWs_.ptr, bsScanWs_.bytes))
                    return e;
                prevLayout = gLayout_.as<uint32_t>();
            }
            // when the previous sort's largest bucket was well below the guard, this one's is not read first: it
            // reaches pinned word kMaxWord with the global update's synchronization and guards the next sort
            const bool defer = lastMaxBucket_ <= kMaxBucket / 4;
            uint32_t largest = ~0u;
            uint32_t* maxOut = defer ? pin + kMaxWord : &largest;
            *maxOut          = ~0u;
            if (auto e = sortByLeafBuckets<K, T>(s, keys_.as<K>(), n, gTree_[gCur_].template as<K>(), numGlobal_, x, y,
                                                 z, h, xyzh_[0].as<T>(), xyzh_[1].as<T>(), xyzh_[2].as<T>(),
                                                 xyzh_[3].as<T>(), perm_.as<uint32_t>(), keyBuf_.as<K>(),
                                                 valBuf_.as<uint32_t>(), bsWs_, bsScanWs_, &bucketed, prevLayout,
                                                 &hintWs_, maxOut, defer))
                return e;
            maxPending_    = defer;
            lastMaxBucket_ = defer ? ~0u : largest;
        }
        warmPos_ = bucketed; // the bucket offsets are the old global boundaries' positions for the first update
        if (bucketed) return HIPCSTONE_OK;
        return A::sortGather(s, keys_.as<K>(), perm_.as<uint32_t>(), n, x, y, z, h, xyzh_[0].as<T>(), xyzh_[1].as<T>(),
                             xyzh_[2].as<T>(), xyzh_[3].as<T>(), keyBuf_.as<K>(), valBuf_.as<uint32_t>(),
                             sortWs_.ptr, sortWs_.bytes);
    }

    hipcstone_status reserveParticles(size_t n)
    {
        if (n <= capN_) return HIPCSTONE_OK;
        for (auto* b : {&keys_, &keyBuf_})
            if (auto e = b->reserve(n * sizeof(K))) return e;
        for (auto* b : {&perm_, &valBuf_})
            if (auto e = b->reserve(n * sizeof(uint32_t))) return e;
        for (int a = 0; a < 4; ++a)
            if (auto e = xyzh_[a].reserve(n * sizeof(T))) return e;
        if (auto e = sortWs_.reserve(A::sortWs(n))) return e;
        if (auto e = mmPartials_.reserve(size_t(kMmBlocks) * 6 * sizeof(T))) return e;
        if (auto e = mmOut_.reserve(6 * sizeof(T))) return e;
        if (auto e = small_.reserve(64)) return e;
        capN_ = n;
        return HIPCSTONE_OK;
    }

    // global leaf array buffers for cap leaves, keeping the current tree
    hipcstone_status reserveGlobal(int32_t cap)
    {
        if (cap <= capGlobal_) return HIPCSTONE_OK;
        std::vector<K> keepTree;
        std::vector<uint32_t> keepCounts;
        if (numGlobal_ > 0)
        {
            keepTree.resize(numGlobal_ + 1);
            keepCounts.resize(numGlobal_);
            HC_CHECK(hipMemcpy(keepTree.data(), gTree_[gCur_].ptr, keepTree.size() * sizeof(K), hipMemcpyDeviceToHost));
            HC_CHECK(hipMemcpy(keepCounts.data(), gCounts_[gCur_].ptr, keepCounts.size() * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
        }
        int32_t c = std::max(cap, capGlobal_ * 2);
        for (int t = 0; t < 2; ++t)
        {
            if (auto e = gTree_[t].reserve((size_t(c) + 1) * sizeof(K))) return e;
            if (auto e = gCounts_[t].reserve(size_t(c) * sizeof(uint32_t))) return e;
        }
        if (auto e = updWs_.reserve(hipcstone_update_octree_workspace_size(c))) return e;
        if (numGlobal_ > 0)
        {
            HC_CHECK(hipMemcpy(gTree_[gCur_].ptr, keepTree.data(), keepTree.size() * sizeof(K), hipMemcpyHostToDevice));
            HC_CHECK(hipMemcpy(gCounts_[gCur_].ptr, keepCounts.data(), keepCounts.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice));
        }
        capGlobal_ = c;
        return HIPCSTONE_OK;
    }

    // one global step (tree/update_mpi.hpp:64-98 at one rank): 0 if the input tree was converged, else the largest
    // leaf count of the new tree
    hipcstone_status globalStep(hipcstone_stream s, uint32_t* result)
    {
        // the first update after a bucket sort through the same global leaves takes its node counts from the
        // bucket offsets
        const bool warm = warmPos_;
        warmPos_        = false;
        for (int attempt = 0; attempt < 8; ++attempt)
        {
            int32_t nn = numGlobal_;
            int other  = 1 - gCur_;
            int conv   = 0;
            hipcstone_status e;
            if (warm)
            {
                if (auto g = posWs_.reserve((size_t(capGlobal_) + 1) * sizeof(uint32_t))) return g;
                e = updateOctreeWarm<K>(s, keys_.as<K>(), n_, bucket_, gTree_[gCur_].template as<K>(),
                                        gCounts_[gCur_].template as<uint32_t>(), &nn, gTree_[other].template as<K>(),
                                        gCounts_[other].template as<uint32_t>(), capGlobal_, updWs_.ptr,
                                        updWs_.bytes, &conv, bsWs_.as<uint32_t>() + numGlobal_,
                                        posWs_.as<uint32_t>());
            }
            else
            {
                e = A::update(s, keys_.as<K>(), n_, bucket_, gTree_[gCur_].template as<K>(),
                              gCounts_[gCur_].template as<uint32_t>(), &nn, gTree_[other].template as<K>(),
                              gCounts_[other].template as<uint32_t>(), capGlobal_, updWs_.ptr, updWs_.bytes, &conv);
            }
            if (e == HIPCSTONE_CAPACITY)
            {
                if (auto g = reserveGlobal(nn)) return g;
                continue;
            }
            if (e) return e;
            gCur_      = other;
            numGlobal_ = nn;
            if (conv)
            {
                *result = 0;
                return HIPCSTONE_OK;
            }
            uint32_t* dmax = small_.as<uint32_t>();
            setU32Kernel<<<1, 1, 0, stream(s)>>>(dmax, 0u);
            HC_CHECK_LAUNCH("setU32Kernel");
            unsigned blocks = std::min<unsigned>(1024, ceilDiv(size_t(nn), 256));
            maxCountKernel<<<blocks, 256, 0, stream(s)>>>(gCounts_[gCur_].template as<uint32_t>(), nn, dmax);
            HC_CHECK_LAUNCH("maxCountKernel");
            HC_CHECK(hipMemcpyAsync(result, dmax, sizeof(uint32_t), hipMemcpyDeviceToHost, stream(s)));
            HC_CHECK(hipStreamSynchronize(stream(s)));
            return HIPCSTONE_OK;
        }
        return invalid("domain: global leaf capacity growth did not converge");
    }

    // domain/assignment.hpp:115-123
    hipcstone_status globalUpdate(hipcstone_stream s)
    {
        if (first_)
        {
            if (auto e = reserveGlobal(4096)) return e;
            // the spanning tree of initialDomainSplits at one rank is the root; counts bucketSize - 1
            setKeysKernel<K><<<1, 1, 0, stream(s)>>>(gTree_[gCur_].template as<K>(), K(0), nodeRange<K>(0));
            HC_CHECK_LAUNCH("setKeysKernel");
            setU32Kernel<<<1, 1, 0, stream(s)>>>(gCounts_[gCur_].template as<uint32_t>(), bucket_ - 1);
            HC_CHECK_LAUNCH("setU32Kernel");
            numGlobal_ = 1;
        }
        uint32_t m  = 0;
        globalSteps_ = 1;
        if (auto e = globalStep(s, &m)) return e;
        if (first_ || m >= 8u * bucket_)
        {
            while (true)
            {
                ++globalSteps_;
                if (auto e = globalStep(s, &m)) return e;
                if (m <= bucket_) break;
                if (globalSteps_ > 128) return invalid("domain: global tree update did not settle in 128 steps");
            }
        }
        return HIPCSTONE_OK;
    }

    // buffers sized by the linked focus tree of L leaves (none of them holds data that outlives a relink)
    hipcstone_status reserveFocus(int32_t L)
    {
        size_t Lc = size_t(L), Nc = size_t(L) + size_t(L - 1) / 7;
        if (auto e = fLeafCounts_.reserve(Lc * sizeof(uint32_t))) return e;
        if (auto e = leafOps_.reserve((Lc + 1) * sizeof(int32_t))) return e;
        if (auto e = layout_.reserve((Lc + 1) * sizeof(uint32_t))) return e;
        if (auto e = scanWs_.reserve(std::max(scan::workspaceBytes<int32_t>(Lc + 1), scan::workspaceBytes<uint32_t>(Lc + 1))))
            return e;
        if (auto e = buildWs_.reserve(A::buildWs(L))) return e;
        if (auto e = prefixes_.reserve(Nc * sizeof(K))) return e;
        for (auto* b : {&childOffsets_, &internalToLeaf_, &leafToInternal_, &ops_, &prot_})
            if (auto e = b->reserve(Nc * sizeof(int32_t))) return e;
        if (auto e = parents_.reserve(std::max<size_t>(1, (Nc - 1) / 8 + 1) * sizeof(int32_t))) return e;
        if (auto e = levelRange_.reserve(32 * sizeof(int32_t))) return e;
        if (auto e = fNodeCounts_.reserve(Nc * sizeof(uint32_t))) return e;
        if (auto e = centers_.reserve(Nc * 3 * sizeof(T))) return e;
        if (auto e = sizes_.reserve(Nc * 3 * sizeof(T))) return e;
        return HIPCSTONE_OK;
    }

    hipcstone_status linkFocus(hipcstone_stream s)
    {
        if (auto e = reserveFocus(fL_)) return e;
        fN_ = fL_ + (fL_ - 1) / 7;
        return A::build(s, fLeaves_.as<K>(), fL_, prefixes_.as<K>(), childOffsets_.as<int32_t>(),
                        parents_.as<int32_t>(), levelRange_.as<int32_t>(), internalToLeaf_.as<int32_t>(),
                        leafToInternal_.as<int32_t>(), buildWs_.ptr, buildWs_.bytes, nullptr);
    }

    // FocusedOctree construction (octree_focus_mpi.hpp:61-87): root leaf with count bucketSizeFocus + 1
    hipcstone_status initFocus(hipcstone_stream s)
    {
        fL_ = 1;
        if (auto e = fLeaves_.reserve(2 * sizeof(K))) return e;
        setKeysKernel<K><<<1, 1, 0, stream(s)>>>(fLeaves_.as<K>(), K(0), nodeRange<K>(0));
        HC_CHECK_LAUNCH("setKeysKernel");
        if (auto e = linkFocus(s)) return e;
        setU32Kernel<<<1, 1, 0, stream(s)>>>(fNodeCounts_.as<uint32_t>(), bucketFocus_ + 1);
        HC_CHECK_LAUNCH("setU32Kernel");
        setU32Kernel<<<1, 1, 0, stream(s)>>>(fLeafCounts_.as<uint32_t>(), bucketFocus_ + 1);
        HC_CHECK_LAUNCH("setU32Kernel");
        return HIPCSTONE_OK;
    }

    // CombinedUpdate::updateFocus (octree_focus.hpp:64-111) with focus [0, nodeRange(0)) and the global leaves as
    // mandatory keys; relinks the new tree
    hipcstone_status focusUpdate(hipcstone_stream s, int* converged)
    {
        const int32_t L = fL_, N = fN_, NI = fN_ - fL_;
        const int32_t M = numGlobal_ + 1;
        int32_t* flags  = small_.as<int32_t>(); // [0] status, [1] some protected op != 1, [2] some leaf op != 1
        zeroI32Kernel<<<1, 64, 0, stream(s)>>>(flags, 3);
        HC_CHECK_LAUNCH("zeroI32Kernel");
        focusNodeOpsKernel<K><<<ceilDiv(size_t(N), 256), 256, 0, stream(s)>>>(
            prefixes_.as<K>(), childOffsets_.as<int32_t>(), parents_.as<int32_t>(), fNodeCounts_.as<uint32_t>(),
            nullptr, N, K(0), nodeRange<K>(0), bucketFocus_, ops_.as<int32_t>());
        HC_CHECK_LAUNCH("focusNodeOpsKernel");
        enforceKeysKernel<K><<<ceilDiv(size_t(M), 256), 256, 0, stream(s)>>>(
            gTree_[gCur_].template as<K>(), M, prefixes_.as<K>(), childOffsets_.as<int32_t>(),
            parents_.as<int32_t>(), ops_.as<int32_t>(), flags);
        HC_CHECK_LAUNCH("enforceKeysKernel");
        protectAncestorsKernel<K><<<ceilDiv(size_t(N), 256), 256, 0, stream(s)>>>(
            prefixes_.as<K>(), parents_.as<int32_t>(), ops_.as<int32_t>(), N, prot_.as<int32_t>(), flags + 1);
        HC_CHECK_LAUNCH("protectAncestorsKernel");
        gatherLeafOpsKernel<<<ceilDiv(size_t(L) + 1, 256), 256, 0, stream(s)>>>(
            prot_.as<int32_t>(), leafToInternal_.as<int32_t>() + NI, L, leafOps_.as<int32_t>(), flags + 2);
        HC_CHECK_LAUNCH("gatherLeafOpsKernel");
        if (auto e = scan::exclusiveScan<int32_t>(stream(s), leafOps_.as<int32_t>(), leafOps_.as<int32_t>(),
                                                  size_t(L) + 1, 0, nullptr, scanWs_.ptr, scanWs_.bytes))
            return e;
        int32_t hostFlags[4];
        HC_CHECK(hipMemcpyAsync(hostFlags, flags, 3 * sizeof(int32_t), hipMemcpyDeviceToHost, stream(s)));
        HC_CHECK(hipMemcpyAsync(hostFlags + 3, leafOps_.as<int32_t>() + L, sizeof(int32_t), hipMemcpyDeviceToHost,
                                stream(s)));
        // the device's fault word rides along with this synchronization (guards of the sort and earlier kernels)
        uint32_t faults = 0;
        if (auto e = enqueueFaultRead(stream(s), &faults)) return e;
        HC_CHECK(hipStreamSynchronize(stream(s)));
        if (auto e = checkFaults(faults)) return e;
        const int status = hostFlags[0];
        const int32_t newL = hostFlags[3];
        int conv = hostFlags[1] == 0;
        if (status == 1) conv = hostFlags[2] == 0;
        else if (status >= 2) conv = 0;

        if (newL <= 0) return invalid("domain: focus rebalance produced no leaves");
        if (auto e = reserveFocusNew(newL + M + 1)) return e;
        if (auto e = A::rebalance(s, fLeaves_.as<K>(), L, newL, leafOps_.as<int32_t>(), fLeavesNew_.as<K>())) return e;
        int32_t numKeys = newL + 1;
        if (status == 3)
        {
            // injectKeys (focus/inject.hpp:21-45): the mandatory keys {0, nodeRange(0), global leaves} are sorted,
            // unique and a cornerstone array, so their spanning tree is themselves; merge, sort, drop duplicates
            HC_CHECK(hipMemcpyAsync(fLeavesNew_.as<K>() + numKeys, gTree_[gCur_].ptr, size_t(M) * sizeof(K),
                                    hipMemcpyDeviceToDevice, stream(s)));
            int32_t total = numKeys + M;
            if (auto e = A::sortKeys(s, fLeavesNew_.as<K>(), size_t(total), injKeyBuf_.as<K>(), injSortWs_.ptr,
                                     injSortWs_.bytes))
                return e;
            uniqueFlagsKernel<K><<<ceilDiv(size_t(total), 256), 256, 0, stream(s)>>>(fLeavesNew_.as<K>(), total,
                                                                                   injFlags_.as<uint32_t>());
            HC_CHECK_LAUNCH("uniqueFlagsKernel");
            if (auto e = scan::exclusiveScan<uint32_t>(stream(s), injFlags_.as<uint32_t>(), injFlags_.as<uint32_t>(),
                                                       size_t(total), 0u, injFlags_.as<uint32_t>() + total,
                                                       injScanWs_.ptr, injScanWs_.bytes))
                return e;
            uniqueScatterKernel<K><<<ceilDiv(size_t(total), 256), 256, 0, stream(s)>>>(
                fLeavesNew_.as<K>(), injFlags_.as<uint32_t>(), total, injKeyBuf_.as<K>());
            HC_CHECK_LAUNCH("uniqueScatterKernel");
            uint32_t unique = 0;
            HC_CHECK(hipMemcpyAsync(&unique, injFlags_.as<uint32_t>() + total, sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, stream(s)));
            HC_CHECK(hipStreamSynchronize(stream(s)));
            numKeys = int32_t(unique);
            HC_CHECK(hipMemcpyAsync(fLeavesNew_.ptr, injKeyBuf_.ptr, size_t(numKeys) * sizeof(K),
                                    hipMemcpyDeviceToDevice, stream(s)));
        }
        swap(fLeaves_, fLeavesNew_);
        fL_ = numKeys - 1;
        if (auto e = linkFocus(s)) return e;
        *converged = conv;
        return HIPCSTONE_OK;
    }

    // destination buffers of the rebalanced (and possibly injected) leaf array of up to `keys` keys
    hipcstone_status reserveFocusNew(int32_t keys)
    {
        size_t c = size_t(keys);
        if (auto e = fLeavesNew_.reserve(c * sizeof(K))) return e;
        if (auto e = injKeyBuf_.reserve(c * sizeof(K))) return e;
        if (auto e = injFlags_.reserve((c + 1) * sizeof(uint32_t))) return e;
        if (auto e = injSortWs_.reserve(A::keySortWs(c))) return e;
        if (auto e = injScanWs_.reserve(scan::workspaceBytes<uint32_t>(c))) return e;
        return HIPCSTONE_OK;
    }

    // focus counts of the linked tree (octree_focus_mpi.hpp:233-245 at one rank)
    hipcstone_status focusCounts(hipcstone_stream s)
    {
        nodeKeyCountsKernel<K><<<ceilDiv(size_t(fN_), 256), 256, 0, stream(s)>>>(
            prefixes_.as<K>(), fN_, keys_.as<K>(), n_, fNodeCounts_.as<uint32_t>());
        HC_CHECK_LAUNCH("nodeKeyCountsKernel");
        gatherU32Kernel<<<ceilDiv(size_t(fL_), 256), 256, 0, stream(s)>>>(
            fNodeCounts_.as<uint32_t>(), leafToInternal_.as<int32_t>() + (fN_ - fL_), fL_, fLeafCounts_.as<uint32_t>());
        HC_CHECK_LAUNCH("gatherU32Kernel");
        return HIPCSTONE_OK;
    }

    // geometric centers and the particle layout of the focus tree; particles back into the caller's arrays
    hipcstone_status finish(hipcstone_stream s, T* x, T* y, T* z, T* h, void* keysOut)
    {
        if (auto e = A::centers(s, prefixes_.as<K>(), fN_, centers_.as<T>(), sizes_.as<T>(), &box_)) return e;
        if (auto e = scan::exclusiveScan<uint32_t>(stream(s), fLeafCounts_.as<uint32_t>(), layout_.as<uint32_t>(),
                                                   size_t(fL_), 0u, layout_.as<uint32_t>() + fL_, scanWs_.ptr,
                                                   scanWs_.bytes))
            return e;
        T* dst[4] = {x, y, z, h};
        for (int a = 0; a < 4; ++a)
            HC_CHECK(hipMemcpyAsync(dst[a], xyzh_[a].ptr, n_ * sizeof(T), hipMemcpyDeviceToDevice, stream(s)));
        if (keysOut) HC_CHECK(hipMemcpyAsync(keysOut, keys_.ptr, n_ * sizeof(K), hipMemcpyDeviceToDevice, stream(s)));
        return HIPCSTONE_OK;
    }

    uint32_t bucket_, bucketFocus_;
    float invTheta_;
    bool centersCurrent_ = false;
    T lo_[3]{}, hi_[3]{};
    uint8_t bnd_[3]{};
    Box box_{};
    bool first_       = true;
    size_t n_         = 0;
    size_t capN_      = 0;
    int32_t numGlobal_ = 0, capGlobal_ = 0, gCur_ = 0;
    int32_t fL_ = 0, fN_ = 0;
    int globalSteps_ = 0, focusSteps_ = 0, focusConverged_ = 0;

    DeviceBuf keys_, keyBuf_, perm_, valBuf_, sortWs_, mmPartials_, mmOut_, small_, bsWs_, bsScanWs_, gLayout_,
        hintWs_;
    // pinned word of the calling thread that receives a deferred largest bucket (the pipeline's is word 9)
    static constexpr int kMaxWord = 10;
    uint32_t lastMaxBucket_ = ~0u;  // largest bucket of the last bucket sort, ~0 when unknown
    bool maxPending_        = false; // the last sort's largest bucket is on its way to pinned word kMaxWord
    DeviceBuf xyzh_[4];
    DeviceBuf gTree_[2], gCounts_[2], updWs_, posWs_;
    bool warmPos_ = false; // this call's bucket sort left the old global boundaries' positions in bsWs_
    DeviceBuf fLeaves_, fLeavesNew_, fLeafCounts_, fNodeCounts_, leafOps_, ops_, prot_;
    DeviceBuf prefixes_, childOffsets_, parents_, levelRange_, internalToLeaf_, leafToInternal_, buildWs_;
    DeviceBuf centers_, sizes_, layout_, scanWs_;
    DeviceBuf injKeyBuf_, injFlags_, injSortWs_, injScanWs_, propScratch_;
    DeviceBuf expCenters_;
};

} // namespace hipcstone

struct hipcstone_domain
{
    std::unique_ptr<hipcstone::DomainBase> impl;
};

using namespace hipcstone;

extern "C"
{

hipcstone_status hipcstone_domain_create(hipcstone_domain** d, int key_bits, int real_bits, uint32_t bucket_size,
                                         uint32_t bucket_size_focus, float theta, const void* box)
{
    if (!d || !box) return invalid("domain_create: null pointer");
    if (!(theta > 0.f)) return invalid("domain_create: theta must be positive");
    if (bucket_size == 0 || bucket_size_focus == 0) return invalid("domain_create: bucket sizes must be positive");
    if (bucket_size_focus > bucket_size) return invalid("domain_create: bucket_size_focus above bucket_size");
    auto* h = new hipcstone_domain;
    if (key_bits == 64 && real_bits == 32)
        h->impl = std::make_unique<Domain<uint64_t, float>>(bucket_size, bucket_size_focus, theta,
                                                           *static_cast<const hipcstone_box_f32*>(box));
    else if (key_bits == 64 && real_bits == 64)
        h->impl = std::make_unique<Domain<uint64_t, double>>(bucket_size, bucket_size_focus, theta,
                                                            *static_cast<const hipcstone_box_f64*>(box));
    else if (key_bits == 32 && real_bits == 32)
        h->impl = std::make_unique<Domain<uint32_t, float>>(bucket_size, bucket_size_focus, theta,
                                                           *static_cast<const hipcstone_box_f32*>(box));
    else if (key_bits == 32 && real_bits == 64)
        h->impl = std::make_unique<Domain<uint32_t, double>>(bucket_size, bucket_size_focus, theta,
                                                            *static_cast<const hipcstone_box_f64*>(box));
    else
    {
        delete h;
        return invalid("domain_create: key_bits must be 32|64 and real_bits 32|64");
    }
    *d = h;
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_domain_destroy(hipcstone_domain* d)
{
    delete d;
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_domain_sync(hipcstone_domain* d, hipcstone_stream s, void* x, void* y, void* z, void* h,
                                       size_t n, void* keys_out)
{
    if (!d) return invalid("domain_sync: null handle");
    return d->impl->sync(s, x, y, z, h, n, keys_out);
}

hipcstone_status hipcstone_domain_get_view(const hipcstone_domain* d, hipcstone_domain_view* v)
{
    if (!d || !v) return invalid("domain_get_view: null pointer");
    d->impl->view(v);
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_domain_gather_properties(hipcstone_domain* d, hipcstone_stream s, void* const* arrays,
                                                    const size_t* elem_bytes, int num_arrays)
{
    if (!d) return invalid("domain_gather_properties: null handle");
    return d->impl->gatherProperties(s, arrays, elem_bytes, num_arrays);
}

hipcstone_status hipcstone_domain_sync_grav(hipcstone_domain* d, hipcstone_stream s, void* x, void* y, void* z,
                                            void* h, void* m, size_t n, void* keys_out)
{
    if (!d) return invalid("domain_sync_grav: null handle");
    return d->impl->syncGrav(s, x, y, z, h, m, n, keys_out);
}

hipcstone_status hipcstone_domain_update_expansion_centers(hipcstone_domain* d, hipcstone_stream s, const void* x,
                                                           const void* y, const void* z, const void* m)
{
    if (!d) return invalid("domain_update_expansion_centers: null handle");
    return d->impl->updateExpansionCenters(s, x, y, z, m);
}

hipcstone_status hipcstone_domain_set_timing(hipcstone_domain* d, int enable)
{
    if (!d) return invalid("domain_set_timing: null handle");
    d->impl->timer.enable(enable != 0);
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_domain_stage_times(hipcstone_domain* d, double* ms, int32_t* launches)
{
    if (!d) return invalid("domain_stage_times: null handle");
    return d->impl->timer.collect(ms, launches);
}

} // extern "C"
