// Batched clouds: many independent particle clouds through keys -> sort (+ gather) -> leaf array -> link -> centers
// -> layout -> neighbors, every stage over all clouds in one launch, and the warm drift steps enqueued without any
// host synchronization so that a sequence of them can be captured in one HIP graph.
//
// Each cloud c (at most 512) is exactly what a plain pipeline gives for it alone with 32-bit keys (Morton or
// Hilbert), its own box and its own bucket size: keys, order, leaf array, counts, layout, neighbor rows and counts are
// bit-identical (checked against one plain pipeline per cloud). The design:
//
//  * Composite keys. Particle keys are the 64-bit (c << 54) | (k30 << 24), with k30 the cloud's own 30-bit key of the
//    particle (sfc/sfc.hpp:142-179 with the cloud's box). Every 64-bit octree level below level 3 then corresponds to
//    one level of the cloud's own 32-bit tree (node range 2^(54-3l) = 2^(30-3(l-3)) * 2^24), and the 512 level-3
//    nodes are the clouds. One stable radix sort of the key bits [24, 63) sorts every cloud at once (the segmented
//    sort), and the clouds stay contiguous in cloud order.
//  * One cornerstone leaf array over all clouds. Its leaves at level >= 3 inside cloud c are the cloud's leaves
//    (shifted); the rebalance rules of tree/csarray.hpp:250-290 apply per cloud with the cloud's bucket size and the
//    cloud's deepest level (3 + 10), and no merge ever forms a node above a cloud (level < 3), so one warm
//    updateOctree step over the whole array is one step of every cloud (csarray.hpp:408-426).
//  * Device-side sizes. The leaf count after a rebalance lives on the device; every kernel after it runs a fixed grid
//    that loops over the device count (the arrays have a fixed capacity; a rebalance that would exceed it raises
//    HIPCSTONE_FAULT_CAPACITY and keeps every write in bounds). A warm step therefore never synchronizes the host.
//  * A link without a sort. The reference orders nodes level by level (a sort of all prefixes, tree/octree.hpp:167-196).
//    The neighbor walk only needs the 8 children of a node to be consecutive, in child order: here internal node P
//    gets the slot s(P) that the unsorted layout of tree/octree.hpp:77-100 gives it ((tid + binaryKeyWeight) / 7, a
//    bijection onto [0, numInternal)), the root's slot is swapped with 0, and the children of P are nodes
//    1 + 8 s(P) + digit. Each node finds its parent's slot with one binary search of the leaf array. The node order
//    differs from the reference's, but each cloud's walk visits the same leaves in the same order, so neighbor rows
//    are identical.
//  * Geometry per cloud: a node at level >= 3 is decoded as the cloud's own 32-bit key (as a Hilbert key, like
//    nodeFpCenters, focus/source_center.hpp:145-158) in the cloud's box.
//  * The neighbor search is the plain kernel in its segmented mode: waves never straddle clouds, and each starts at
//    its cloud's root node.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "common.hpp"
#include "runtime.hpp"
#include "host_util.hpp"
#include "scan.hpp"
#include "sort.hpp"
#include "synth.hpp"

namespace hipcstone
{

template<class T>
hipcstone_status findNeighborsSegmented(hipStream_t s, const T* x, const T* y, const T* z, const T* h,
                                        const typename BoxOf<T>::type& box, const int32_t* childOffsets,
                                        const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                        const T* sizes, int32_t numNodesCap, uint32_t ngmax, uint32_t* neighbors,
                                        uint32_t* neighborsCount, const uint32_t* waves, uint32_t numWaves,
                                        const int32_t* roots, const nbwalk::ChildGroup* groups);

namespace batched
{

using K = uint64_t;
using T = float;
using Box = BoxOf<float>::type;

constexpr int kMaxClouds        = 512;
constexpr unsigned kCloudLevels = 3;
constexpr int kCloudShift       = 54;
constexpr int kLocalShift       = 24;
constexpr unsigned kLocalLevels = 10; // 32-bit keys
constexpr unsigned kMaxLevelB   = kCloudLevels + kLocalLevels;
constexpr int kGrid             = 2048; // blocks of the grid-stride kernels over device-side counts
constexpr int kThreads          = 256;

__device__ inline uint32_t cloudOfKey(K key) { return uint32_t(key >> kCloudShift); }

// the cloud holding particle i: the last c with offsets[c] <= i (offsets staged in LDS, numClouds + 1 entries)
__device__ inline uint32_t cloudOfParticle(const uint32_t* shOff, uint32_t numClouds, uint32_t i)
{
    uint32_t lo = 0, hi = numClouds; // answer in [lo, hi)
    while (hi - lo > 1)
    {
        uint32_t mid = (lo + hi) / 2;
        if (shOff[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline void stageOffsets(uint32_t* shOff, const uint32_t* offsets, uint32_t numClouds)
{
    for (uint32_t c = threadIdx.x; c <= numClouds; c += blockDim.x)
        shOff[c] = offsets[c];
    __syncthreads();
}

// ---- drift: particle i of cloud c moves as the plain displacement of that cloud alone moves its local particle
// i - offsets[c] under seed + c (synth.hpp)
__global__ __launch_bounds__(kThreads) void driftKernel(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z,
                                                        const T* __restrict__ h, size_t n,
                                                        const uint32_t* __restrict__ offsets, uint32_t numClouds,
                                                        const Box* __restrict__ boxes, T scale, uint64_t seed,
                                                        uint64_t step)
{
    __shared__ uint32_t shOff[kMaxClouds + 1];
    stageOffsets(shOff, offsets, numClouds);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        uint32_t c = cloudOfParticle(shOff, numClouds, uint32_t(i));
        displaceAt(x, y, z, h, i, uint64_t(i - shOff[c]), scale, seed + c, step, boxes[c]);
    }
}

// ---- composite keys
template<int Curve>
__global__ __launch_bounds__(kThreads) void keysKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                       const T* __restrict__ z, size_t n,
                                                       const uint32_t* __restrict__ offsets, uint32_t numClouds,
                                                       const Box* __restrict__ boxes, K* __restrict__ keys)
{
    __shared__ uint32_t shOff[kMaxClouds + 1];
    stageOffsets(shOff, offsets, numClouds);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        uint32_t c   = cloudOfParticle(shOff, numClouds, uint32_t(i));
        uint32_t k30 = sfc3D<uint32_t, Curve>(x[i], y[i], z[i], boxes[c]);
        keys[i]      = (K(c) << kCloudShift) | (K(k30) << kLocalShift);
    }
}

// ---- full build (first call): the unique converged tree of every cloud, built top down from the cloud nodes
// (the fixpoint argument of the plain pipeline's computeOctree, per cloud: a node is internal iff it holds more
// than the cloud's bucket size and lies above the cloud's deepest level)
struct SplitNode
{
    K start;
    uint32_t lo, hi;
};

__device__ inline uint32_t lowerBoundKeys(const K* __restrict__ keys, uint32_t lo, uint32_t hi, K v)
{
    while (lo < hi)
    {
        uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// cloud slots: cloud c < numClouds is a frontier node holding its particles, a slot without a cloud a leaf of 0
__global__ void cloudNodesKernel(const uint32_t* __restrict__ offsets, uint32_t numClouds,
                                 const uint32_t* __restrict__ bucket, SplitNode* __restrict__ frontier,
                                 K* __restrict__ leafKeys, uint32_t* __restrict__ leafCounts,
                                 uint32_t* __restrict__ counters)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= uint32_t(kMaxClouds)) return;
    K start = K(c) << kCloudShift;
    uint32_t lo = c < numClouds ? offsets[c] : 0u, hi = c < numClouds ? offsets[c + 1] : 0u;
    if (hi - lo > bucket[c]) // above the bucket size: internal (level 3 < the deepest level)
    {
        uint32_t at   = atomicAdd(counters, 1u);
        frontier[at] = SplitNode{start, lo, hi};
    }
    else
    {
        uint32_t at    = atomicAdd(counters + 1, 1u);
        leafKeys[at]   = start;
        leafCounts[at] = hi - lo;
    }
}

__global__ __launch_bounds__(256) void splitLevelKernel(const K* __restrict__ keys, const SplitNode* __restrict__ in,
                                                        uint32_t nIn, unsigned childLevel,
                                                        const uint32_t* __restrict__ bucket,
                                                        SplitNode* __restrict__ out, K* __restrict__ leafKeys,
                                                        uint32_t* __restrict__ leafCounts, uint32_t* counters)
{
    const uint32_t gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t node = gid >> 3;
    const unsigned k    = gid & 7u;
    const bool valid    = node < nIn;
    SplitNode f{K(0), 0u, 0u};
    if (valid) f = in[node];
    const K start = f.start + K(k) * nodeRange<K>(childLevel);
    uint32_t b    = (valid && k > 0) ? lowerBoundKeys(keys, f.lo, f.hi, start) : f.lo;
    uint32_t e    = __shfl_down(b, 1, 8);
    if (k == 7) e = f.hi;
    const uint32_t cnt  = e - b;
    const bool internal = valid && cnt > bucket[cloudOfKey(f.start)] && childLevel < kMaxLevelB;
    const bool leaf     = valid && !internal;
    const uint64_t mi   = ballot(internal);
    const uint64_t ml   = ballot(leaf);
    uint32_t bi = 0, bl = 0;
    if (laneId() == 0)
    {
        if (mi) bi = atomicAdd(counters, uint32_t(__builtin_popcountll(mi)));
        if (ml) bl = atomicAdd(counters + 1, uint32_t(__builtin_popcountll(ml)));
    }
    bi = uniform(bi);
    bl = uniform(bl);
    if (internal) out[bi + laneRank(mi)] = SplitNode{start, b, e};
    if (leaf)
    {
        leafKeys[bl + laneRank(ml)]   = start;
        leafCounts[bl + laneRank(ml)] = cnt;
    }
}

__global__ void setPairKernel(uint32_t* p, uint32_t a, uint32_t b)
{
    p[0] = a;
    p[1] = b;
}

// ---- warm update (tree/csarray.hpp:250-389 per cloud), over the device-side leaf count
// ops[i] for i < L; ops[L] = 0 and ops[(L, cap]] = 0 so that a scan of cap + 1 entries gives the new count at ops[L]
__global__ __launch_bounds__(kThreads) void nodeOpsKernel(const K* __restrict__ tree, const int32_t* __restrict__ dL,
                                                          const uint32_t* __restrict__ counts,
                                                          const uint32_t* __restrict__ bucket,
                                                          int32_t* __restrict__ ops, int32_t cap)
{
    const int32_t L = *dL;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= cap; i += gridDim.x * blockDim.x)
    {
        if (i >= L)
        {
            ops[i] = 0;
            continue;
        }
        K thisNode     = tree[i];
        unsigned level = treeLevel(K(tree[i + 1] - thisNode));
        int op         = 1;
        if (level >= kCloudLevels) // nodes above a cloud are never leaves of a converged tree; they stay as they are
        {
            const uint32_t b = bucket[cloudOfKey(thisNode)];
            bool merged      = false;
            // merging 8 siblings forms their parent, which must lie inside the cloud (level - 1 >= 3)
            if (level > kCloudLevels)
            {
                int siblingIdx = int(octalDigit(thisNode, level));
                bool siblings  = tree[i - siblingIdx + 8] == tree[i - siblingIdx] + nodeRange<K>(level - 1);
                if (siblings && siblingIdx > 0)
                {
                    const uint32_t* g = counts + i - siblingIdx;
                    uint64_t parent   = uint64_t(g[0]) + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7];
                    if (parent <= uint64_t(b))
                    {
                        op     = 0;
                        merged = true;
                    }
                }
            }
            if (!merged)
            {
                uint32_t c = counts[i];
                if (c > b * 512 && level + 3 < kMaxLevelB) op = 4096;
                else if (c > b * 64 && level + 2 < kMaxLevelB) op = 512;
                else if (c > b * 8 && level + 1 < kMaxLevelB) op = 64;
                else if (c > b && level < kMaxLevelB) op = 8;
            }
        }
        ops[i] = op;
    }
}

// tree/csarray.hpp:339-366 processNode; the new count is ops[L] (after the scan). A new count above the capacity
// raises the capacity fault and copies the old tree instead, so that the step goes on over a valid (unchanged) leaf
// array and every later index stays in bounds.
__global__ __launch_bounds__(kThreads) void rebalanceKernel(const K* __restrict__ tree, const int32_t* __restrict__ dL,
                                                            const int32_t* __restrict__ ops, K* __restrict__ newTree,
                                                            int32_t cap, uint32_t* __restrict__ faults)
{
    const int32_t L    = *dL;
    const int32_t newL = ops[L];
    if (newL > cap)
    {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(faults, HIPCSTONE_FAULT_CAPACITY);
        for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= L; i += gridDim.x * blockDim.x)
            newTree[i] = tree[i];
        return;
    }
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= L; i += gridDim.x * blockDim.x)
    {
        if (i == L)
        {
            newTree[newL] = tree[L];
            continue;
        }
        K thisNode     = tree[i];
        unsigned level = treeLevel(K(tree[i + 1] - thisNode));
        int32_t op     = ops[i + 1] - ops[i];
        int32_t dst    = ops[i];
        if (op == 1) { newTree[dst] = thisNode; }
        else if (op >= 8)
        {
            unsigned levelDiff = op == 8 ? 1u : log8ceil(K(uint32_t(op)));
            K step             = nodeRange<K>(level + levelDiff);
            for (int32_t s = 0; s < op; ++s)
                newTree[dst + s] = thisNode + K(s) * step;
        }
    }
}

// the new leaf count becomes current (the old one if the rebalance did not fit)
__global__ void commitCountKernel(int32_t* dL, const int32_t* __restrict__ ops, int32_t cap)
{
    int32_t newL = ops[*dL];
    if (newL <= cap) *dL = newL;
}

// counts of the leaves [0, L) (exact, as computeNodeCounts gives them) and zeros in [L, cap]: each leaf's range of
// the sorted keys by two binary searches inside its cloud's particle range
__global__ __launch_bounds__(kThreads) void countsKernel(const K* __restrict__ tree, const int32_t* __restrict__ dL,
                                                         const K* __restrict__ keys,
                                                         const uint32_t* __restrict__ offsets, uint32_t numClouds,
                                                         uint32_t* __restrict__ counts, int32_t cap)
{
    const int32_t L = *dL;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= cap; i += gridDim.x * blockDim.x)
    {
        if (i >= L)
        {
            counts[i] = 0;
            continue;
        }
        K a = tree[i], b = tree[i + 1];
        uint32_t c = cloudOfKey(a);
        uint32_t lo = c < numClouds ? offsets[c] : 0u, hi = c < numClouds ? offsets[c + 1] : 0u;
        counts[i] = lowerBoundKeys(keys, lo, hi, b) - lowerBoundKeys(keys, lo, hi, a);
    }
}

// ---- link
__device__ inline int32_t lowerBoundLeaves(const K* __restrict__ leaves, int32_t n, K v)
{
    int32_t lo = 0, hi = n;
    while (lo < hi)
    {
        int32_t mid = lo + (hi - lo) / 2;
        if (leaves[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// slot (octIndex, tree/octree.hpp:86-99) of internal node (key, level >= 0): created at the boundary between the
// last leaf of its child 3 and the first of its child 4
__device__ inline int32_t internalSlot(const K* __restrict__ leaves, int32_t L, K key, unsigned level)
{
    K boundary  = key + K(4) * nodeRange<K>(level + 1);
    int32_t tid = lowerBoundLeaves(leaves, L + 1, boundary) - 1;
    tid         = tid < 0 ? 0 : tid; // only for an invalid array (leaves[0] is 0 < boundary)
    return (tid + binaryKeyWeight(leaves[tid], level)) / 7;
}

__device__ inline int32_t remapSlot(int32_t slot, int32_t rootSlot)
{
    return slot == rootSlot ? 0 : (slot == 0 ? rootSlot : slot);
}

// node index of (key, level): the root is 0, else 1 + 8 * slot(parent) + digit
__device__ inline int32_t nodeIndex(const K* __restrict__ leaves, int32_t L, int32_t rootSlot, K key, unsigned level)
{
    if (level == 0) return 0;
    K parentKey = key & ~(nodeRange<K>(level - 1) - 1);
    return 1 + 8 * remapSlot(internalSlot(leaves, L, parentKey, level - 1), rootSlot) +
           int32_t(octalDigit(key, level));
}

// meta[0] = root slot, meta[1] = number of nodes, meta[2] = number of internal nodes
__global__ void linkMetaKernel(const K* __restrict__ leaves, const int32_t* __restrict__ dL, int32_t* __restrict__ meta)
{
    const int32_t L = *dL;
    const int32_t I = (L - 1) / 7;
    meta[0]         = L > 1 ? internalSlot(leaves, L, K(0), 0) : 0;
    meta[1]         = L + I;
    meta[2]         = I;
}

__global__ __launch_bounds__(kThreads) void linkNodesKernel(const K* __restrict__ leaves, const int32_t* __restrict__ dL,
                                                            const int32_t* __restrict__ meta, K* __restrict__ prefixes,
                                                            int32_t* __restrict__ childOffsets,
                                                            int32_t* __restrict__ parents,
                                                            int32_t* __restrict__ internalToLeaf,
                                                            int32_t* __restrict__ leafToInternal)
{
    const int32_t L        = *dL;
    const int32_t rootSlot = meta[0];
    const int32_t I        = meta[2];
    const int32_t N        = meta[1];
    // a valid cornerstone array maps every node into [0, N) (a bijection); anything else is dropped
    auto inTree = [&](int32_t idx) { return idx >= 0 && idx < N; };
    for (int32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < L; t += gridDim.x * blockDim.x)
    {
        // leaf t, and the internal node its right boundary creates
        {
            K key          = leaves[t];
            unsigned level = treeLevel(K(leaves[t + 1] - key));
            int32_t idx    = L == 1 ? 0 : nodeIndex(leaves, L, rootSlot, key, level);
            if (!inTree(idx)) continue;
            prefixes[idx]       = encodePlaceholderBit(key, int(3 * level));
            childOffsets[idx]   = 0;
            internalToLeaf[idx] = t;
            leafToInternal[I + t] = idx;
            // the internal node created between leaves t and t + 1 (tree/octree.hpp:90-98)
            if (t < L - 1)
            {
                unsigned pl = unsigned(commonPrefix(key, leaves[t + 1]));
                if (pl % 3 == 0)
                {
                    unsigned lv = pl / 3;
                    int32_t o   = (t + binaryKeyWeight(key, lv)) / 7;
                    K nodeKey   = lv == 0 ? K(0) : (key & ~(nodeRange<K>(lv) - 1));
                    int32_t ix  = nodeIndex(leaves, L, rootSlot, nodeKey, lv);
                    int32_t s   = remapSlot(o, rootSlot);
                    if (!inTree(ix) || o < 0 || o >= I || s < 0 || s >= I) continue;
                    prefixes[ix]       = encodePlaceholderBit(key, int(pl));
                    childOffsets[ix]   = 1 + 8 * s;
                    internalToLeaf[ix] = -1 - o;
                    leafToInternal[o]  = ix;
                    parents[s]         = ix;
                }
            }
        }
    }
}

// node centers and half sizes: nodes at level >= 3 in their cloud's key space and box (Hilbert decode, as
// nodeFpCenters); the nodes above the clouds get zeros (no walk visits them)
__global__ __launch_bounds__(kThreads) void centersKernel(const K* __restrict__ prefixes, const int32_t* __restrict__ meta,
                                                          const Box* __restrict__ boxes, uint32_t numClouds,
                                                          T* __restrict__ centers, T* __restrict__ sizes)
{
    const int32_t N = meta[1];
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
    {
        K prefix       = prefixes[i];
        K startKey     = decodePlaceholderBit(prefix);
        unsigned level = decodePrefixLength(prefix) / 3;
        size_t o       = size_t(i) * 3;
        uint32_t c     = cloudOfKey(startKey);
        if (level < kCloudLevels || c >= numClouds)
        {
            for (int d = 0; d < 3; ++d)
                centers[o + d] = sizes[o + d] = T(0);
            continue;
        }
        uint32_t local = uint32_t(startKey >> kLocalShift) & ((1u << 30) - 1u);
        int ib[6];
        hilbertIBox(local, level - kCloudLevels, ib);
        const Box& box = boxes[c];
        constexpr T uL = T(1.) / T(1u << kLocalLevels);
        T hx           = T(0.5) * uL * box.len[0];
        T hy           = T(0.5) * uL * box.len[1];
        T hz           = T(0.5) * uL * box.len[2];
        centers[o + 0] = box.lim[0] + T(ib[1] + ib[0]) * hx;
        centers[o + 1] = box.lim[2] + T(ib[3] + ib[2]) * hy;
        centers[o + 2] = box.lim[4] + T(ib[5] + ib[4]) * hz;
        sizes[o + 0]   = T(ib[1] - ib[0]) * hx;
        sizes[o + 1]   = T(ib[3] - ib[2]) * hy;
        sizes[o + 2]   = T(ib[5] - ib[4]) * hz;
    }
}

// the root node of every cloud (its level-3 node)
__global__ void cloudRootsKernel(const K* __restrict__ leaves, const int32_t* __restrict__ dL,
                                 const int32_t* __restrict__ meta, uint32_t numClouds, int32_t* __restrict__ roots)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numClouds) return;
    roots[c] = nodeIndex(leaves, *dL, meta[0], K(c) << kCloudShift, kCloudLevels);
}

__global__ void setI32Kernel(int32_t* p, int32_t v) { *p = v; }

// ---------------------------------------------------------------------------------------------------------------------
class Batch
{
public:
    hipcstone_status init(int curve, uint32_t numClouds, const uint64_t* counts, const Box* boxes,
                          const uint32_t* buckets, uint32_t ngmax, double leafCapacityFactor)
    {
        if (numClouds == 0 || numClouds > uint32_t(kMaxClouds)) return invalid("batched: 1 to 512 clouds");
        if (curve != HIPCSTONE_HILBERT && curve != HIPCSTONE_MORTON) return invalid("batched: curve");
        curve_ = curve;
        C_     = numClouds;
        ngmax_ = ngmax;
        offsets_.assign(C_ + 1, 0);
        for (uint32_t c = 0; c < C_; ++c)
        {
            if (buckets[c] == 0) return invalid("batched: bucket size 0");
            offsets_[c + 1] = offsets_[c] + counts[c];
        }
        if (offsets_[C_] == 0 || offsets_[C_] > (uint64_t(1) << 30) - 1)
            return invalid("batched: total particle count must lie in [1, 2^30)");
        n_ = size_t(offsets_[C_]);
        std::vector<uint32_t> off32(C_ + 1), bucket(kMaxClouds, 1u << 30);
        for (uint32_t c = 0; c <= C_; ++c)
            off32[c] = uint32_t(offsets_[c]);
        for (uint32_t c = 0; c < C_; ++c)
            bucket[c] = buckets[c];
        hostBoxes_.assign(boxes, boxes + C_);
        buckets_.assign(buckets, buckets + C_);
        // the wave table: waves of at most 64 targets that never straddle a cloud
        std::vector<uint32_t> waves;
        for (uint32_t c = 0; c < C_; ++c)
            for (uint32_t f = off32[c]; f < off32[c + 1]; f += kWave)
            {
                waves.push_back(f);
                waves.push_back(std::min(f + uint32_t(kWave), off32[c + 1]));
                waves.push_back(c);
            }
        numWaves_ = uint32_t(waves.size() / 3);

        // leaf capacity: a multiple of the particles over the smallest bucket size (warm steps change the leaf
        // count by a few per cent; the first build reports a larger need and regrows before any step is captured)
        uint32_t minBucket = *std::min_element(buckets, buckets + C_);
        capL_ = int32_t(std::min<double>(double(INT32_MAX / 2), leafCapacityFactor * double(n_) / minBucket +
                                                                    8 * kMaxClouds + 4096));
        if (auto e = allocParticles()) return e;
        if (auto e = allocTree(capL_)) return e;
        HC_CHECK(hipMemcpy(offsetsD_.ptr, off32.data(), off32.size() * 4, hipMemcpyHostToDevice));
        HC_CHECK(hipMemcpy(bucketD_.ptr, bucket.data(), bucket.size() * 4, hipMemcpyHostToDevice));
        HC_CHECK(hipMemcpy(boxesD_.ptr, boxes, C_ * sizeof(Box), hipMemcpyHostToDevice));
        HC_CHECK(hipMemcpy(wavesD_.ptr, waves.data(), waves.size() * 4, hipMemcpyHostToDevice));
        return HIPCSTONE_OK;
    }

    // first call: keys, full sort, converged trees (host-driven level loop), then the rest of a step
    hipcstone_status build(hipcstone_stream s)
    {
        hipStream_t st = stream(s);
        if (auto e = keysAndSort(st, xyzh_[cur_][0].as<T>(), xyzh_[cur_][1].as<T>(), xyzh_[cur_][2].as<T>(),
                                 xyzh_[cur_][3].as<T>()))
            return e;
        if (auto e = fullTree(st)) return e;
        if (auto e = afterTree(st)) return e;
        built_ = true;
        return HIPCSTONE_OK;
    }

    // one warm step from the current particle set: optional drift in place, keys, sort, one update, link, centers,
    // layout, neighbors. Enqueues only (no synchronization, no allocation): capturable.
    hipcstone_status step(hipcstone_stream s, float driftScale, uint64_t seed, uint64_t stepIndex)
    {
        if (!built_) return invalid("batched step: build first");
        hipStream_t st = stream(s);
        T* x = xyzh_[cur_][0].as<T>();
        T* y = xyzh_[cur_][1].as<T>();
        T* z = xyzh_[cur_][2].as<T>();
        T* h = xyzh_[cur_][3].as<T>();
        if (driftScale != 0.0f)
        {
            driftKernel<<<kGrid, kThreads, 0, st>>>(x, y, z, h, n_, offsetsD_.as<uint32_t>(), C_,
                                                    boxesD_.as<Box>(), T(driftScale), seed, stepIndex);
            HC_CHECK_LAUNCH("batched driftKernel");
        }
        if (auto e = keysAndSort(st, x, y, z, h)) return e;
        if (auto e = warmTree(st)) return e;
        return afterTree(st);
    }

    // upload a fresh particle set (x,y,z,h of all clouds, in cloud order) as the next build's or step's input
    hipcstone_status setParticles(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, int device)
    {
        T* dst[4]       = {xyzh_[cur_][0].as<T>(), xyzh_[cur_][1].as<T>(), xyzh_[cur_][2].as<T>(),
                           xyzh_[cur_][3].as<T>()};
        const T* src[4] = {x, y, z, h};
        for (int a = 0; a < 4; ++a)
            HC_CHECK(hipMemcpyAsync(dst[a], src[a], n_ * sizeof(T),
                                    device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s)));
        return HIPCSTONE_OK;
    }

    hipcstone_status view(hipcstone_stream s, hipcstone_batched_view* v)
    {
        int32_t meta[3] = {0, 0, 0};
        int32_t L       = 0;
        HC_CHECK(hipMemcpyAsync(meta, metaD_.ptr, sizeof(meta), hipMemcpyDeviceToHost, stream(s)));
        HC_CHECK(hipMemcpyAsync(&L, dL_.ptr, sizeof(L), hipMemcpyDeviceToHost, stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s)));
        v->num_clouds       = C_;
        v->num_particles    = n_;
        v->num_leaves       = L;
        v->num_nodes        = meta[1];
        v->leaf_capacity    = capL_;
        v->num_waves        = numWaves_;
        v->keys             = keys_.ptr;
        v->perm             = perm_.ptr;
        v->x                = xyzh_[cur_][0].ptr;
        v->y                = xyzh_[cur_][1].ptr;
        v->z                = xyzh_[cur_][2].ptr;
        v->h                = xyzh_[cur_][3].ptr;
        v->leaves           = tree_[treeCur_].ptr;
        v->counts           = counts_.ptr;
        v->layout           = layout_.ptr;
        v->prefixes         = prefixes_.ptr;
        v->child_offsets    = childOffsets_.ptr;
        v->internal_to_leaf = internalToLeaf_.ptr;
        v->centers          = centers_.ptr;
        v->sizes            = sizes_.ptr;
        v->cloud_roots      = roots_.ptr;
        v->neighbors        = nb_.ptr;
        v->neighbor_counts  = nc_.ptr;
        return HIPCSTONE_OK;
    }

    size_t deviceBytes() const { return allocated_; }

    // capture `steps` warm steps (step indices first .. first + steps - 1) on s into one executable graph; nothing
    // runs until replay. The step count must be even: every step swaps the particle and leaf buffer pairs, so an even
    // count leaves the host's view of the buffers where a replay leaves the data.
    hipcstone_status capture(hipcstone_stream s, int steps, float driftScale, uint64_t seed, uint64_t first)
    {
        if (!built_) return invalid("batched capture: build first");
        if (steps <= 0 || steps % 2 != 0) return invalid("batched capture: the step count must be even and positive");
        hipStream_t st = stream(s);
        if (!st) return invalid("batched capture: needs a created stream (the null stream cannot be captured)");
        if (!faultWord()) return HIPCSTONE_HIP_ERROR; // resolved before the capture begins
        dropGraph();
        HC_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipcstone_status err = HIPCSTONE_OK;
        for (int k = 0; k < steps && err == HIPCSTONE_OK; ++k)
            err = step(s, driftScale, seed, first + uint64_t(k));
        hipGraph_t g      = nullptr;
        hipError_t endErr = hipStreamEndCapture(st, &g);
        if (err != HIPCSTONE_OK)
        {
            if (g) (void)hipGraphDestroy(g);
            return err;
        }
        if (endErr != hipSuccess) return hipFail(endErr, "hipStreamEndCapture");
        hipError_t ie = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess)
        {
            graph_ = nullptr;
            return hipFail(ie, "hipGraphInstantiate");
        }
        return HIPCSTONE_OK;
    }

    hipcstone_status replay(hipcstone_stream s)
    {
        if (!graph_) return invalid("batched replay: nothing captured");
        HC_CHECK(hipGraphLaunch(graph_, stream(s)));
        return HIPCSTONE_OK;
    }

    ~Batch() { dropGraph(); }

private:
    void dropGraph()
    {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }

    hipcstone_status reserve(DeviceBuf& b, size_t bytes)
    {
        size_t before = b.bytes;
        if (auto e = b.reserve(bytes)) return e;
        allocated_ += b.bytes - before;
        return HIPCSTONE_OK;
    }

    hipcstone_status allocParticles()
    {
        for (auto* b : {&keys_, &keyBuf_})
            if (auto e = reserve(*b, n_ * sizeof(K))) return e;
        for (auto* b : {&perm_, &valBuf_, &nc_})
            if (auto e = reserve(*b, n_ * sizeof(uint32_t))) return e;
        for (int s = 0; s < 2; ++s)
            for (int a = 0; a < 4; ++a)
                if (auto e = reserve(xyzh_[s][a], n_ * sizeof(T))) return e;
        if (auto e = reserve(sortWs_, sortGatherWorkspaceBytes<K>(n_))) return e;
        if (auto e = reserve(nb_, std::max<size_t>(1, n_ * size_t(ngmax_)) * sizeof(uint32_t))) return e;
        if (auto e = reserve(offsetsD_, (kMaxClouds + 1) * sizeof(uint32_t))) return e;
        if (auto e = reserve(bucketD_, kMaxClouds * sizeof(uint32_t))) return e;
        if (auto e = reserve(boxesD_, kMaxClouds * sizeof(Box))) return e;
        if (auto e = reserve(wavesD_, size_t(numWaves_) * 3 * sizeof(uint32_t))) return e;
        if (auto e = reserve(roots_, kMaxClouds * sizeof(int32_t))) return e;
        if (auto e = reserve(dL_, 256)) return e;
        if (auto e = reserve(metaD_, 256)) return e;
        if (auto e = reserve(splitCounters_, 256)) return e;
        return HIPCSTONE_OK;
    }

    hipcstone_status allocTree(int32_t cap)
    {
        size_t L1 = size_t(cap) + 1;
        size_t N  = size_t(cap) + size_t(cap) / 7 + 1;
        for (int t = 0; t < 2; ++t)
            if (auto e = reserve(tree_[t], L1 * sizeof(K))) return e;
        if (auto e = reserve(counts_, L1 * sizeof(uint32_t))) return e;
        if (auto e = reserve(ops_, L1 * sizeof(int32_t))) return e;
        if (auto e = reserve(layout_, L1 * sizeof(uint32_t))) return e;
        if (auto e = reserve(scanWs_, scan::workspaceBytes<uint32_t>(L1))) return e;
        if (auto e = reserve(opsScanWs_, scan::workspaceBytes<int32_t>(L1))) return e;
        if (auto e = reserve(prefixes_, N * sizeof(K))) return e;
        for (auto* b : {&childOffsets_, &internalToLeaf_, &leafToInternal_, &parents_})
            if (auto e = reserve(*b, N * sizeof(int32_t))) return e;
        for (auto* b : {&centers_, &sizes_})
            if (auto e = reserve(*b, N * 3 * sizeof(T))) return e;
        if (auto e = reserve(groups_, std::max<size_t>(nbwalk::childGroupCount(int32_t(N)), 1) *
                                          sizeof(nbwalk::ChildGroup)))
            return e;
        capL_ = cap;
        capN_ = int32_t(N);
        return HIPCSTONE_OK;
    }

    // keys of the current particles, stable sort of bits [24, 63) with x,y,z,h gathered into the other set
    hipcstone_status keysAndSort(hipStream_t st, const T* x, const T* y, const T* z, const T* h)
    {
        int out = 1 - cur_;
        if (curve_ == HIPCSTONE_HILBERT)
            keysKernel<HIPCSTONE_HILBERT><<<kGrid, kThreads, 0, st>>>(x, y, z, n_, offsetsD_.as<uint32_t>(), C_,
                                                                     boxesD_.as<Box>(), keys_.as<K>());
        else
            keysKernel<HIPCSTONE_MORTON><<<kGrid, kThreads, 0, st>>>(x, y, z, n_, offsetsD_.as<uint32_t>(), C_,
                                                                    boxesD_.as<Box>(), keys_.as<K>());
        HC_CHECK_LAUNCH("batched keysKernel");
        if (auto e = sortGatherBits<K, T>(st, keys_.as<K>(), perm_.as<uint32_t>(), n_, x, y, z, h,
                                          xyzh_[out][0].as<T>(), xyzh_[out][1].as<T>(), xyzh_[out][2].as<T>(),
                                          xyzh_[out][3].as<T>(), keyBuf_.as<K>(), valBuf_.as<uint32_t>(),
                                          sortWs_.ptr, sortWs_.bytes, kLocalShift, kCloudShift + 9))
            return e;
        cur_ = out;
        return HIPCSTONE_OK;
    }

    // the converged trees of all clouds (first call), host-driven level by level
    hipcstone_status fullTree(hipStream_t st)
    {
        // frontiers: every frontier node holds more particles than its bucket size, so a level has fewer than n
        if (auto e = reserve(frontier_[0], std::max<size_t>(kMaxClouds, n_) * sizeof(SplitNode))) return e;
        if (auto e = reserve(frontier_[1], std::max<size_t>(kMaxClouds, n_) * sizeof(SplitNode))) return e;
        setPairKernel<<<1, 1, 0, st>>>(splitCounters_.as<uint32_t>(), 0u, 0u);
        HC_CHECK_LAUNCH("setPairKernel");
        // leaves go straight into tree_[1 - treeCur_] / counts_ (unsorted), grown on demand below
        size_t leafStage = std::max<size_t>(size_t(capL_) + 1, kMaxClouds);
        if (auto e = reserve(splitKeys_, leafStage * sizeof(K))) return e;
        if (auto e = reserve(splitCounts_, leafStage * sizeof(uint32_t))) return e;
        cloudNodesKernel<<<kMaxClouds / 64, 64, 0, st>>>(offsetsD_.as<uint32_t>(), C_, bucketD_.as<uint32_t>(),
                                                         frontier_[0].as<SplitNode>(), splitKeys_.as<K>(),
                                                         splitCounts_.as<uint32_t>(), splitCounters_.as<uint32_t>());
        HC_CHECK_LAUNCH("cloudNodesKernel");
        uint32_t host[2];
        HC_CHECK(hipMemcpyAsync(host, splitCounters_.ptr, sizeof(host), hipMemcpyDeviceToHost, st));
        HC_CHECK(hipStreamSynchronize(st));
        uint32_t nIn = host[0], leaves = host[1];
        int cur = 0;
        for (unsigned childLevel = kCloudLevels + 1; nIn > 0; ++childLevel)
        {
            if (childLevel > kMaxLevelB) return invalid("batched: split below the deepest level");
            const size_t nOutMax = size_t(nIn) * 8;
            if (nOutMax > frontier_[1 - cur].bytes / sizeof(SplitNode))
            {
                DeviceBuf nb;
                if (auto e = nb.reserve(nOutMax * sizeof(SplitNode))) return e;
                swap(nb, frontier_[1 - cur]);
            }
            if (leaves + nOutMax > splitKeys_.bytes / sizeof(K))
            {
                if (auto e = growKeep(splitKeys_, (leaves + nOutMax) * sizeof(K), leaves * sizeof(K), st)) return e;
                if (auto e = growKeep(splitCounts_, (leaves + nOutMax) * sizeof(uint32_t), leaves * 4, st)) return e;
            }
            setPairKernel<<<1, 1, 0, st>>>(splitCounters_.as<uint32_t>(), 0u, leaves);
            HC_CHECK_LAUNCH("setPairKernel");
            splitLevelKernel<<<ceilDiv(nOutMax, 256), 256, 0, st>>>(
                keys_.as<K>(), frontier_[cur].as<SplitNode>(), nIn, childLevel, bucketD_.as<uint32_t>(),
                frontier_[1 - cur].as<SplitNode>(), splitKeys_.as<K>(), splitCounts_.as<uint32_t>(),
                splitCounters_.as<uint32_t>());
            HC_CHECK_LAUNCH("batched splitLevelKernel");
            HC_CHECK(hipMemcpyAsync(host, splitCounters_.ptr, sizeof(host), hipMemcpyDeviceToHost, st));
            HC_CHECK(hipStreamSynchronize(st));
            nIn    = host[0];
            leaves = host[1];
            cur    = 1 - cur;
        }
        const int32_t L = int32_t(leaves);
        // capacity: the first build fixes it for the captured steps (twice the converged count plus a margin)
        int32_t want = std::max<int32_t>(capL_, int32_t(std::min<int64_t>(INT32_MAX / 2, 2 * int64_t(L) + 4096)));
        if (want > capL_)
        {
            if (auto e = allocTree(want)) return e;
        }
        K* tree          = tree_[treeCur_].as<K>();
        uint32_t* counts = counts_.as<uint32_t>();
        HC_CHECK(hipMemcpyAsync(tree, splitKeys_.ptr, size_t(L) * sizeof(K), hipMemcpyDeviceToDevice, st));
        HC_CHECK(hipMemcpyAsync(counts, splitCounts_.ptr, size_t(L) * 4, hipMemcpyDeviceToDevice, st));
        // leaf keys are distinct; sort them (with their counts) in the split buffers' space
        if (auto e = sortKeysCounts(st, tree, counts, L)) return e;
        const K endKey = nodeRange<K>(0);
        HC_CHECK(hipMemcpyAsync(tree + L, &endKey, sizeof(K), hipMemcpyHostToDevice, st));
        setI32Kernel<<<1, 1, 0, st>>>(dL_.as<int32_t>(), L);
        HC_CHECK_LAUNCH("setI32Kernel");
        // zero the counts past L (the layout scan runs over the capacity)
        countsTail(st);
        HC_CHECK(hipStreamSynchronize(st)); // endKey lives on this stack frame
        return HIPCSTONE_OK;
    }

    hipcstone_status sortKeysCounts(hipStream_t st, K* tree, uint32_t* counts, int32_t L);

    void countsTail(hipStream_t st)
    {
        countsKernel<<<kGrid, kThreads, 0, st>>>(tree_[treeCur_].as<K>(), dL_.as<int32_t>(), keys_.as<K>(),
                                                 offsetsD_.as<uint32_t>(), C_, counts_.as<uint32_t>(), capL_);
    }

    // one warm updateOctree step of every cloud over the device-side leaf count
    hipcstone_status warmTree(hipStream_t st)
    {
        uint32_t* faults = faultWord();
        if (!faults) return HIPCSTONE_HIP_ERROR;
        int other = 1 - treeCur_;
        nodeOpsKernel<<<kGrid, kThreads, 0, st>>>(tree_[treeCur_].as<K>(), dL_.as<int32_t>(), counts_.as<uint32_t>(),
                                                  bucketD_.as<uint32_t>(), ops_.as<int32_t>(), capL_);
        HC_CHECK_LAUNCH("batched nodeOpsKernel");
        if (auto e = scan::exclusiveScan<int32_t>(st, ops_.as<int32_t>(), ops_.as<int32_t>(), size_t(capL_) + 1, 0,
                                                  nullptr, opsScanWs_.ptr, opsScanWs_.bytes))
            return e;
        rebalanceKernel<<<kGrid, kThreads, 0, st>>>(tree_[treeCur_].as<K>(), dL_.as<int32_t>(), ops_.as<int32_t>(),
                                                    tree_[other].as<K>(), capL_, faults);
        HC_CHECK_LAUNCH("batched rebalanceKernel");
        commitCountKernel<<<1, 1, 0, st>>>(dL_.as<int32_t>(), ops_.as<int32_t>(), capL_);
        HC_CHECK_LAUNCH("batched commitCountKernel");
        treeCur_ = other;
        countsKernel<<<kGrid, kThreads, 0, st>>>(tree_[treeCur_].as<K>(), dL_.as<int32_t>(), keys_.as<K>(),
                                                 offsetsD_.as<uint32_t>(), C_, counts_.as<uint32_t>(), capL_);
        HC_CHECK_LAUNCH("batched countsKernel");
        return HIPCSTONE_OK;
    }

    // link, centers, layout, cloud roots, neighbors
    hipcstone_status afterTree(hipStream_t st)
    {
        const K* leaves = tree_[treeCur_].as<K>();
        linkMetaKernel<<<1, 1, 0, st>>>(leaves, dL_.as<int32_t>(), metaD_.as<int32_t>());
        HC_CHECK_LAUNCH("batched linkMetaKernel");
        linkNodesKernel<<<kGrid, kThreads, 0, st>>>(leaves, dL_.as<int32_t>(), metaD_.as<int32_t>(), prefixes_.as<K>(),
                                                    childOffsets_.as<int32_t>(), parents_.as<int32_t>(),
                                                    internalToLeaf_.as<int32_t>(), leafToInternal_.as<int32_t>());
        HC_CHECK_LAUNCH("batched linkNodesKernel");
        centersKernel<<<kGrid, kThreads, 0, st>>>(prefixes_.as<K>(), metaD_.as<int32_t>(), boxesD_.as<Box>(), C_,
                                                  centers_.as<T>(), sizes_.as<T>());
        HC_CHECK_LAUNCH("batched centersKernel");
        cloudRootsKernel<<<ceilDiv(C_, 64), 64, 0, st>>>(leaves, dL_.as<int32_t>(), metaD_.as<int32_t>(), C_,
                                                          roots_.as<int32_t>());
        HC_CHECK_LAUNCH("batched cloudRootsKernel");
        if (auto e = scan::exclusiveScan<uint32_t>(st, counts_.as<uint32_t>(), layout_.as<uint32_t>(), size_t(capL_),
                                                   0u, layout_.as<uint32_t>() + capL_, scanWs_.ptr, scanWs_.bytes))
            return e;
        // the compact child blocks of all clouds' trees (child blocks start at 1 + 8 s): blocks past the live
        // node count are built from stale arrays and never read
        const uint32_t numGroups = uint32_t(nbwalk::childGroupCount(capN_));
        if (numGroups > 0)
        {
            nbwalk::childGroupsKernel<><<<ceilDiv(size_t(numGroups), 256), 256, 0, st>>>(
                centers_.as<T>(), sizes_.as<T>(), childOffsets_.as<int32_t>(), internalToLeaf_.as<int32_t>(),
                numGroups, groups_.as<nbwalk::ChildGroup>());
            HC_CHECK_LAUNCH("batched childGroupsKernel");
        }
        Box any = hostBoxes_[0];
        for (int d = 0; d < 3; ++d)
            any.bnd[d] = HIPCSTONE_OPEN;
        return findNeighborsSegmented<T>(st, xyzh_[cur_][0].as<T>(), xyzh_[cur_][1].as<T>(), xyzh_[cur_][2].as<T>(),
                                         xyzh_[cur_][3].as<T>(), any, childOffsets_.as<int32_t>(),
                                         internalToLeaf_.as<int32_t>(), layout_.as<uint32_t>(), centers_.as<T>(),
                                         sizes_.as<T>(), capN_, ngmax_, ngmax_ ? nb_.as<uint32_t>() : nullptr,
                                         nc_.as<uint32_t>(), wavesD_.as<uint32_t>(), numWaves_, roots_.as<int32_t>(),
                                         numGroups ? groups_.as<nbwalk::ChildGroup>() : nullptr);
    }

    static hipcstone_status growKeep(DeviceBuf& b, size_t bytes, size_t keep, hipStream_t st)
    {
        if (bytes <= b.bytes) return HIPCSTONE_OK;
        DeviceBuf nb;
        if (auto e = nb.reserve(std::max(bytes, 2 * b.bytes))) return e;
        if (keep) HC_CHECK(hipMemcpyAsync(nb.ptr, b.ptr, keep, hipMemcpyDeviceToDevice, st));
        HC_CHECK(hipStreamSynchronize(st));
        swap(b, nb);
        return HIPCSTONE_OK;
    }

    int curve_        = HIPCSTONE_MORTON;
    uint32_t C_       = 0;
    uint32_t ngmax_   = 0;
    size_t n_         = 0;
    uint32_t numWaves_ = 0;
    int32_t capL_     = 0;
    int32_t capN_     = 0;
    int cur_          = 0;
    int treeCur_      = 0;
    bool built_       = false;
    hipGraphExec_t graph_ = nullptr;
    size_t allocated_ = 0;
    std::vector<uint64_t> offsets_;
    std::vector<Box> hostBoxes_;
    std::vector<uint32_t> buckets_;

    DeviceBuf keys_, keyBuf_, perm_, valBuf_, sortWs_, nb_, nc_;
    DeviceBuf xyzh_[2][4];
    DeviceBuf offsetsD_, bucketD_, boxesD_, wavesD_, roots_, dL_, metaD_, splitCounters_;
    DeviceBuf tree_[2], counts_, ops_, layout_, scanWs_, opsScanWs_;
    DeviceBuf prefixes_, childOffsets_, internalToLeaf_, leafToInternal_, parents_, centers_, sizes_, groups_;
    DeviceBuf frontier_[2], splitKeys_, splitCounts_, leafSortWs_;
};

} // namespace batched
} // namespace hipcstone

extern "C" hipcstone_status hipcstone_sort_by_key_u64(hipcstone_stream, uint64_t*, uint32_t*, size_t, uint64_t*,
                                                      uint32_t*, void*, size_t, int);
extern "C" size_t hipcstone_sort_by_key_workspace_size_u64(size_t n);

namespace hipcstone
{
namespace batched
{

// the converged leaves come out of the level loop unordered; keys are distinct, so a key sort with the counts as
// values orders both (the split buffers hold the sort's double buffers)
hipcstone_status Batch::sortKeysCounts(hipStream_t st, K* tree, uint32_t* counts, int32_t L)
{
    if (auto e = reserve(leafSortWs_, hipcstone_sort_by_key_workspace_size_u64(size_t(L)))) return e;
    return hipcstone_sort_by_key_u64(st, tree, counts, size_t(L), splitKeys_.as<K>(), splitCounts_.as<uint32_t>(),
                                     leafSortWs_.ptr, leafSortWs_.bytes, 0);
}

} // namespace batched
} // namespace hipcstone

using namespace hipcstone;

struct hipcstone_batched
{
    batched::Batch b;
};

extern "C"
{

hipcstone_status hipcstone_batched_create(hipcstone_batched** out, int curve, uint32_t num_clouds,
                                          const uint64_t* counts, const hipcstone_box_f32* boxes,
                                          const uint32_t* bucket_sizes, uint32_t ngmax, double leaf_capacity_factor)
{
    if (!out || !counts || !boxes || !bucket_sizes) return invalid("batched_create: null argument");
    auto p = std::make_unique<hipcstone_batched>();
    if (auto e = p->b.init(curve, num_clouds, counts, boxes, bucket_sizes, ngmax,
                           leaf_capacity_factor > 0 ? leaf_capacity_factor : 0.5))
        return e;
    *out = p.release();
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_batched_destroy(hipcstone_batched* b)
{
    delete b;
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_batched_set_particles(hipcstone_batched* b, hipcstone_stream s, const float* x,
                                                 const float* y, const float* z, const float* h, int device_pointers)
{
    if (!b || !x || !y || !z || !h) return invalid("batched_set_particles: null argument");
    return b->b.setParticles(s, x, y, z, h, device_pointers);
}

hipcstone_status hipcstone_batched_build(hipcstone_batched* b, hipcstone_stream s)
{
    if (!b) return invalid("batched_build: null handle");
    return b->b.build(s);
}

hipcstone_status hipcstone_batched_step(hipcstone_batched* b, hipcstone_stream s, float drift_scale, uint64_t seed,
                                        uint64_t step)
{
    if (!b) return invalid("batched_step: null handle");
    return b->b.step(s, drift_scale, seed, step);
}

hipcstone_status hipcstone_batched_get_view(hipcstone_batched* b, hipcstone_stream s, hipcstone_batched_view* v)
{
    if (!b || !v) return invalid("batched_get_view: null argument");
    return b->b.view(s, v);
}

size_t hipcstone_batched_device_bytes(const hipcstone_batched* b) { return b ? b->b.deviceBytes() : 0; }

hipcstone_status hipcstone_batched_capture(hipcstone_batched* b, hipcstone_stream s, int32_t steps, float drift_scale,
                                           uint64_t seed, uint64_t first_step)
{
    if (!b) return invalid("batched_capture: null handle");
    return b->b.capture(s, steps, drift_scale, seed, first_step);
}

hipcstone_status hipcstone_batched_replay(hipcstone_batched* b, hipcstone_stream s)
{
    if (!b) return invalid("batched_replay: null handle");
    return b->b.replay(s);
}

} // extern "C"
