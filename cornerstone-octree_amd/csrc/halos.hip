// a13/a14: per-leaf search boxes and halo discovery.
//
// Reference semantics: focus/source_center.hpp:28-43 (computeBoundingBox) and traversal/collisions.hpp:25-94
// (findCollisions, findHalos) with traversal/boxoverlap.hpp:127-168 (containedIn) and :280-291 (overlap).
// GPU interface: focus/source_center_gpu.h:40-51, traversal/collisions_gpu.h:46-59.
#include <algorithm>
#include <limits>

#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

// One wave per leaf: lanes take strided particles, then a wave min/max reduction (min/max are exact, so the
// reduction order cannot change the result). The reduction is seeded with the leaf's current center.
template<class T>
__global__ __launch_bounds__(256) void boundingBoxKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                         const T* __restrict__ z, const T* __restrict__ h,
                                                         const uint32_t* __restrict__ layout, int32_t first,
                                                         int32_t last, T scale, T* __restrict__ centers,
                                                         T* __restrict__ sizes)
{
    int32_t leaf = first + int32_t((size_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave);
    if (leaf >= last) return;
    const int lane = laneId();
    T c0 = centers[size_t(leaf) * 3 + 0], c1 = centers[size_t(leaf) * 3 + 1], c2 = centers[size_t(leaf) * 3 + 2];
    T mn0 = c0, mn1 = c1, mn2 = c2, mx0 = c0, mx1 = c1, mx2 = c2;
    uint32_t b = layout[leaf], e = layout[leaf + 1];
    for (uint32_t i = b + lane; i < e; i += kWave)
    {
        T r  = h[i] * scale;
        T p0 = x[i], p1 = y[i], p2 = z[i];
        T l0 = p0 - r, l1 = p1 - r, l2 = p2 - r;
        T u0 = p0 + r, u1 = p1 + r, u2 = p2 + r;
        mn0  = mn0 < l0 ? mn0 : l0;
        mn1  = mn1 < l1 ? mn1 : l1;
        mn2  = mn2 < l2 ? mn2 : l2;
        mx0  = mx0 > u0 ? mx0 : u0;
        mx1  = mx1 > u1 ? mx1 : u1;
        mx2  = mx2 > u2 ? mx2 : u2;
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
    {
        T o;
        o   = __shfl_xor(mn0, off, kWave);
        mn0 = mn0 < o ? mn0 : o;
        o   = __shfl_xor(mn1, off, kWave);
        mn1 = mn1 < o ? mn1 : o;
        o   = __shfl_xor(mn2, off, kWave);
        mn2 = mn2 < o ? mn2 : o;
        o   = __shfl_xor(mx0, off, kWave);
        mx0 = mx0 > o ? mx0 : o;
        o   = __shfl_xor(mx1, off, kWave);
        mx1 = mx1 > o ? mx1 : o;
        o   = __shfl_xor(mx2, off, kWave);
        mx2 = mx2 > o ? mx2 : o;
    }
    if (lane == 0)
    {
        size_t o       = size_t(leaf) * 3;
        centers[o + 0] = (mx0 + mn0) * T(0.5);
        centers[o + 1] = (mx1 + mn1) * T(0.5);
        centers[o + 2] = (mx2 + mn2) * T(0.5);
        sizes[o + 0]   = (mx0 - mn0) * T(0.5);
        sizes[o + 1]   = (mx1 - mn1) * T(0.5);
        sizes[o + 2]   = (mx2 - mn2) * T(0.5);
    }
}

// boxoverlap.hpp:127-152 containedIn(codeStart, codeEnd, center, size, box): corner keys are Hilbert keys
// (SfcKind<KeyType> is HilbertKey, sfc/sfc.hpp:38-40), whatever curve built the tree.
template<class K, class T>
__device__ inline bool boxContainedInRange(K codeStart, K codeEnd, const T c[3], const T s[3],
                                           const typename BoxOf<T>::type& box)
{
    T bmin[3] = {c[0] - s[0], c[1] - s[1], c[2] - s[2]};
    T bmax[3] = {c[0] + s[0], c[1] + s[1], c[2] + s[2]};
    T a0 = bmin[0] - box.lim[0], a1 = bmin[1] - box.lim[2], a2 = bmin[2] - box.lim[4];
    T dFromMin = a0;
    dFromMin   = dFromMin < a1 ? dFromMin : a1;
    dFromMin   = dFromMin < a2 ? dFromMin : a2;
    T b0 = bmax[0] - box.lim[1], b1 = bmax[1] - box.lim[3], b2 = bmax[2] - box.lim[5];
    T dFromMax = b0;
    dFromMax   = dFromMax > b1 ? dFromMax : b1;
    dFromMax   = dFromMax > b2 ? dFromMax : b2;
    if (dFromMin < T(0) || dFromMax > T(0)) return codeStart == 0 && codeEnd == nodeRange<K>(0);

    constexpr unsigned gridDim = 1u << maxLevel<K>();
    T unit                     = T(1) / T(gridDim);
    bmax[0] += box.len[0] * unit;
    bmax[1] += box.len[1] * unit;
    bmax[2] += box.len[2] * unit;
    K lo = sfc3D<K, HIPCSTONE_HILBERT>(bmin[0], bmin[1], bmin[2], box);
    K hi = sfc3D<K, HIPCSTONE_HILBERT>(bmax[0], bmax[1], bmax[2], box);
    K es, ee;
    smallestCommonBox(lo, hi, es, ee);
    return es >= codeStart && ee <= codeEnd;
}

// findHalos (collisions.hpp:25-94): for each local leaf, a walk from the root that calls, at every node it reaches,
// overlaps(node) = !containedIn(node keys, [lowestKey, highestKey)) && overlap(node box, search box), flags the node
// when it returns true and descends into the node's children only then. So node n is flagged iff some active leaf l
// (a local leaf whose search box is not inside the local key range, :88-89) passes overlaps() at n and at every
// ancestor of n. A node with an unflagged parent is never flagged.
//
// GPU design: the tree is walked once, top down, over flagged nodes only, instead of once per leaf. Each flagged
// internal node n carries a witness: an active leaf that passes n and all of n's ancestors. A child c of n is
//   - unflagged if it lies inside the local key range (containedIn holds for c, so for every leaf);
//   - flagged with n's witness if that leaf passes c (its chain to n is already known);
//   - otherwise flagged iff some active leaf passes c and all of c's ancestors: candidates come from a query of an
//     8-ary hierarchy of conservative bounds over the active search boxes, and each candidate's chain is checked
//     with the reference's own test up to the root; the first candidate that passes becomes c's witness.
// Each test is the reference's overlaps() on the reference's operands, so the flag set is the reference's, while the
// work is a few tests per flagged node instead of one walk per leaf (the union walks of leaf groups next to a dense
// cluster re-expand the same subtrees tens of thousands of times).
constexpr int kHaloThreads      = 256;
constexpr unsigned kHaloBlocks  = 2048; // grid of the level kernels (grid-stride over the frontier)
constexpr int kBvhMaxLevels     = 12;   // 8^11 > 2^31 local leaves
constexpr size_t kHaloHeadBytes = 512;  // level starts and counts, root witness

// Conservative bounds of the active search boxes: level 0 holds one box per local leaf (empty for inactive leaves),
// level k+1 the union of 8 consecutive level-k boxes, up to a single box at level top.
struct HaloBvhShape
{
    int32_t offset[kBvhMaxLevels]; // first box of each level
    int32_t count[kBvhMaxLevels];
    int32_t top;
};

struct HaloCounters
{
    uint32_t start[32]; // frontier entries of depth d are front[start[d], start[d] + count[d])
    uint32_t count[32];
    int32_t rootWitness;
    HaloBvhShape bvh;
};
static_assert(sizeof(HaloCounters) <= kHaloHeadBytes, "findHalos counters overlap the frontier");

// Relative slack of the bounds: larger than the rounding error of the reference's overlap() and of the bounds
// themselves, so a leaf that passes overlap() in floating point always lies in bounds that pass the bound test.
template<class T>
__host__ __device__ constexpr T boundsSlack()
{
    return sizeof(T) == 4 ? T(1.0 / (1 << 17)) : T(1.0 / 35184372088832.0); // 2^-17, 2^-45
}

template<class K, class T>
struct HaloArgs
{
    const K* prefixes;
    const int32_t* childOffsets;
    const int32_t* parents;
    const T* nodeCenters;
    const T* nodeSizes;
    const T* searchCenters;
    const T* searchSizes;
    const K* leaves; // the local key range is [leaves[firstNode], leaves[lastNode])
    typename BoxOf<T>::type box;
    int32_t firstNode, lastNode, numNodes;
    uint8_t* flags;
    uint32_t* faults; // HIPCSTONE_FAULT_MALFORMED_TREE when a parent index leaves the tree
};

// containedIn(nodeStart, nodeEnd, lowestKey, highestKey) of the node with this prefix (boxoverlap.hpp:163-168)
template<class K>
__device__ inline bool excludedPrefix(K pfx, K lowestKey, K highestKey)
{
    unsigned pl = decodePrefixLength(pfx);
    K ret       = pfx ^ (K(1) << pl);
    int ns      = 3 * int(maxLevel<K>()) - int(pl);
    K nk1       = ret << ns;
    K nk2       = nk1 + (K(1) << ns);
    return !(nk1 < lowestKey || nk2 > highestKey);
}

// overlap(node box, search box) with the reference's operation order (boxoverlap.hpp:280-291)
template<class T>
__device__ inline bool overlapBox(const T c[3], const T s[3], const T tC[3], const T tS[3],
                                  const typename BoxOf<T>::type& box)
{
    bool ov = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T dx = tC[d] - c[d];
        if (box.bnd[d] == HIPCSTONE_PERIODIC) dx = dx - box.len[d] * rint(dx * box.ilen[d]);
        dx = fabs(dx);
        dx -= s[d];
        dx -= tS[d];
        ov = ov && dx < T(0);
    }
    return ov;
}

template<class T>
__device__ inline void load3(const T* p, int64_t i, T v[3])
{
    v[0] = p[i * 3 + 0];
    v[1] = p[i * 3 + 1];
    v[2] = p[i * 3 + 2];
}

// leaf l's search box passes node a and every ancestor of a
template<class K, class T>
__device__ inline bool chainPasses(const HaloArgs<K, T>& A, int32_t a, const T tC[3], const T tS[3])
{
    for (int step = 0; step < 64; ++step)
    {
        T c[3], s[3];
        load3(A.nodeCenters, a, c);
        load3(A.nodeSizes, a, s);
        if (!overlapBox(c, s, tC, tS, A.box)) return false;
        if (a == 0) return true;
        a = A.parents[(a - 1) / 8];
        if (a < 0 || a >= A.numNodes) break;
    }
    atomicOr(A.faults, HIPCSTONE_FAULT_MALFORMED_TREE); // out-of-range parent, or no root within 64 steps
    return false;
}

// bound test of node box (c, s) against stored bounds; for periodic dimensions the minimum image of the bounds' center
template<class T>
__device__ inline bool boundsOverlap(const T c[3], const T s[3], const T* b, const typename BoxOf<T>::type& box)
{
    bool ov = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T lo = b[d], hi = b[3 + d];
        if (box.bnd[d] == HIPCSTONE_PERIODIC)
        {
            T m  = (lo + hi) * T(0.5);
            T r  = (hi - lo) * T(0.5);
            T dx = c[d] - m;
            dx   = dx - box.len[d] * rint(dx * box.ilen[d]);
            ov   = ov && fabs(dx) < s[d] + r;
        }
        else { ov = ov && c[d] + s[d] > lo && c[d] - s[d] < hi; }
    }
    return ov;
}

// First active leaf (relative index) that passes node c (box cb, sb) and all of c's ancestors starting at parent n,
// or -1. Stackless walk of the implicit 8-ary bounds hierarchy.
template<class K, class T>
__device__ int32_t findWitness(const HaloArgs<K, T>& A, const T* bvh, const HaloBvhShape& H, int32_t n, const T cb[3],
                               const T sb[3])
{
    int k     = H.top;
    int32_t i = 0;
    for (;;)
    {
        if (i < H.count[k] && boundsOverlap(cb, sb, bvh + size_t(H.offset[k] + i) * 6, A.box))
        {
            if (k > 0)
            {
                --k;
                i *= 8;
                continue;
            }
            T tC[3], tS[3];
            int64_t leaf = int64_t(A.firstNode) + i;
            load3(A.searchCenters, leaf, tC);
            load3(A.searchSizes, leaf, tS);
            if (overlapBox(cb, sb, tC, tS, A.box) && chainPasses(A, n, tC, tS)) return i;
        }
        while (k < H.top && ((i & 7) == 7 || i + 1 >= H.count[k]))
        {
            i >>= 3;
            ++k;
        }
        if (k == H.top) return -1;
        ++i;
    }
}

// Level-0 bounds of each local leaf (empty when the leaf is inactive) and the root's witness: the lowest active leaf
// whose search box overlaps the root box.
template<class K, class T>
__global__ __launch_bounds__(256) void haloLeafBoundsKernel(HaloArgs<K, T> A, T* __restrict__ level0,
                                                           HaloCounters* __restrict__ ctr)
{
    const int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= A.lastNode - A.firstNode) return;
    T tC[3], tS[3];
    load3(A.searchCenters, int64_t(A.firstNode) + i, tC);
    load3(A.searchSizes, int64_t(A.firstNode) + i, tS);
    const bool active = !boxContainedInRange<K, T>(A.leaves[A.firstNode], A.leaves[A.lastNode], tC, tS, A.box);
    T* b = level0 + size_t(i) * 6;
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T slack  = (fabs(A.box.lim[2 * d]) + fabs(A.box.lim[2 * d + 1]) + A.box.len[d]) * boundsSlack<T>();
        b[d]     = active ? (tC[d] - tS[d]) - slack : std::numeric_limits<T>::max();
        b[3 + d] = active ? (tC[d] + tS[d]) + slack : std::numeric_limits<T>::lowest();
    }
    if (active)
    {
        T c[3], s[3];
        load3(A.nodeCenters, 0, c);
        load3(A.nodeSizes, 0, s);
        // the witness only decreases: skip the atomic when it already holds a lower leaf
        if (overlapBox(c, s, tC, tS, A.box) &&
            i < __hip_atomic_load(&ctr->rootWitness, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(&ctr->rootWitness, i);
    }
}

template<class T>
__global__ __launch_bounds__(256) void haloBoundsReduceKernel(const T* __restrict__ in, int32_t nIn,
                                                             T* __restrict__ out, int32_t nOut)
{
    const int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= nOut) return;
    T lo[3] = {std::numeric_limits<T>::max(), std::numeric_limits<T>::max(), std::numeric_limits<T>::max()};
    T hi[3] = {std::numeric_limits<T>::lowest(), std::numeric_limits<T>::lowest(), std::numeric_limits<T>::lowest()};
    for (int32_t j = 8 * i; j < 8 * i + 8 && j < nIn; ++j)
    {
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            T a = in[size_t(j) * 6 + d], b = in[size_t(j) * 6 + 3 + d];
            lo[d] = a < lo[d] ? a : lo[d];
            hi[d] = b > hi[d] ? b : hi[d];
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        out[size_t(i) * 6 + d]     = lo[d];
        out[size_t(i) * 6 + 3 + d] = hi[d];
    }
}

__global__ void haloCountersInitKernel(HaloCounters* ctr, HaloBvhShape shape)
{
    int t = threadIdx.x;
    if (t < 32)
    {
        ctr->start[t] = 0;
        ctr->count[t] = 0;
    }
    if (t == 0)
    {
        ctr->rootWitness = INT32_MAX;
#pragma unroll
        for (int k = 0; k < kBvhMaxLevels; ++k)
        {
            ctr->bvh.offset[k] = shape.offset[k];
            ctr->bvh.count[k]  = shape.count[k];
        }
        ctr->bvh.top = shape.top;
    }
}

// The root (collisions.hpp:33-34 via singleTraversal, traversal.hpp:35-45): flagged iff it lies outside the local
// key range and some active leaf overlaps it; if internal, it starts the frontier with its witness.
template<class K, class T>
__global__ void haloRootKernel(HaloArgs<K, T> A, int2* __restrict__ front, HaloCounters* __restrict__ ctr)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int32_t w = ctr->rootWitness;
    if (w == INT32_MAX || excludedPrefix(A.prefixes[0], A.leaves[A.firstNode], A.leaves[A.lastNode])) return;
    A.flags[0]       = 1;
    const int32_t co = A.childOffsets[0];
    if (co > 0 && co + 8 <= A.numNodes)
    {
        front[0]     = make_int2(0, w);
        ctr->count[0] = 1;
    }
}

// Depth d of the walk: tests the 8 children of each frontier node of depth d, flags them, and appends the flagged
// internal children with their witnesses as the frontier of depth d + 1.
template<class K, class T>
__global__ __launch_bounds__(kHaloThreads) void haloLevelKernel(HaloArgs<K, T> A, const T* __restrict__ bvh,
                                                                int2* __restrict__ front, uint32_t capacity,
                                                                HaloCounters* __restrict__ ctr, int depth)
{
    const uint32_t begin = ctr->start[depth];
    const uint32_t n     = begin >= capacity ? 0 : std::min(ctr->count[depth], capacity - begin);
    const uint32_t next  = begin + n; // the next level's entries follow this level's
    const K lowestKey    = A.leaves[A.firstNode];
    const K highestKey   = A.leaves[A.lastNode];
    const HaloBvhShape& H = ctr->bvh;
    if (blockIdx.x == 0 && threadIdx.x == 0) ctr->start[depth + 1] = next;
    const uint64_t total  = uint64_t(n) * 8;
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    const int lane        = laneId();
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x + (threadIdx.x - lane); base < total; base += stride)
    {
        const uint64_t t = base + lane;
        bool emit        = false;
        int32_t c = 0, wit = -1;
        if (t < total)
        {
            const int2 e = front[begin + t / 8];
            c            = A.childOffsets[e.x] + int32_t(t % 8);
            if (!excludedPrefix(A.prefixes[c], lowestKey, highestKey))
            {
                T cb[3], sb[3], tC[3], tS[3];
                load3(A.nodeCenters, c, cb);
                load3(A.nodeSizes, c, sb);
                load3(A.searchCenters, int64_t(A.firstNode) + e.y, tC);
                load3(A.searchSizes, int64_t(A.firstNode) + e.y, tS);
                wit = overlapBox(cb, sb, tC, tS, A.box) ? e.y : findWitness(A, bvh, H, e.x, cb, sb);
                if (wit >= 0)
                {
                    A.flags[c]        = 1;
                    const int32_t cco = A.childOffsets[c];
                    emit              = cco > c && cco + 8 <= A.numNodes;
                }
            }
        }
        const uint64_t eb = ballot(emit);
        if (eb == 0) continue;
        uint32_t pos = 0;
        if (lane == 0) pos = atomicAdd(&ctr->count[depth + 1], uint32_t(__builtin_popcountll(eb)));
        pos            = uniform(pos);
        const uint32_t slot = next + pos + uint32_t(__builtin_popcountll(eb & ((uint64_t(1) << lane) - 1)));
        if (emit && slot < capacity) front[slot] = make_int2(c, wit);
    }
}

template<class T>
static hipcstone_status boundingBox(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h,
                                    const uint32_t* layout, int32_t first, int32_t last, T scale, T* centers,
                                    T* sizes)
{
    if (last <= first) return HIPCSTONE_OK;
    if (!x || !y || !z || !h || !layout || !centers || !sizes) return invalid("computeBoundingBox: null pointer");
    size_t threads = size_t(last - first) * kWave;
    boundingBoxKernel<T><<<ceilDiv(threads, 256), 256, 0, stream(s)>>>(x, y, z, h, layout, first, last, scale,
                                                                         centers, sizes);
    HC_CHECK_LAUNCH("boundingBoxKernel");
    return HIPCSTONE_OK;
}

// Workspace of findHalos: counters, the frontier (at most one entry per node: an internal node enters it at most
// once, from its parent), then the bounds hierarchy.
struct HaloLayout
{
    size_t frontOffset, bvhOffset, bytes;
    int32_t count[kBvhMaxLevels], offset[kBvhMaxLevels], top;
};

inline HaloLayout haloLayout(int32_t numNodes, int32_t numLocal, size_t realBytes)
{
    HaloLayout L{};
    auto align   = [](size_t v) { return (v + 255) / 256 * 256; };
    L.frontOffset = kHaloHeadBytes;
    L.bvhOffset   = align(L.frontOffset + size_t(std::max(numNodes, 1)) * sizeof(int2));
    int32_t n = std::max(numLocal, 1), total = 0, k = 0;
    for (;; ++k)
    {
        L.count[k]  = n;
        L.offset[k] = total;
        total += n;
        if (n == 1 || k + 1 == kBvhMaxLevels) break;
        n = (n + 7) / 8;
    }
    L.top   = k;
    L.bytes = align(L.bvhOffset + size_t(total) * 6 * realBytes);
    return L;
}

template<class K, class T>
static hipcstone_status findHalos(hipcstone_stream s, const K* prefixes, const int32_t* childOffsets,
                                  const int32_t* parents, const T* nodeCenters, const T* nodeSizes, const K* leaves,
                                  const T* searchCenters, const T* searchSizes, const typename BoxOf<T>::type* box,
                                  int32_t firstNode, int32_t lastNode, int32_t numNodes, uint8_t* flags,
                                  void* workspace, size_t workspaceBytes)
{
    if (lastNode <= firstNode) return HIPCSTONE_OK;
    if (!prefixes || !childOffsets || !parents || !nodeCenters || !nodeSizes || !leaves || !searchCenters ||
        !searchSizes || !box || !flags)
        return invalid("findHalos: null pointer");
    if (numNodes <= 0) return invalid("findHalos: numNodes must be positive");
    if (firstNode < 0) return invalid("findHalos: negative first node");
    const HaloLayout lay = haloLayout(numNodes, lastNode - firstNode, sizeof(T));
    void* owned          = nullptr;
    if (!workspace)
    {
        HC_CHECK(hipMallocAsync(&owned, lay.bytes, stream(s)));
        workspace = owned;
    }
    else if (workspaceBytes < lay.bytes)
        return invalid("findHalos: workspace smaller than hipcstone_find_halos_workspace_size()");

    auto* base  = static_cast<char*>(workspace);
    auto* ctr   = reinterpret_cast<HaloCounters*>(base);
    auto* front = reinterpret_cast<int2*>(base + lay.frontOffset);
    T* bvh      = reinterpret_cast<T*>(base + lay.bvhOffset);
    HaloBvhShape shape{};
    for (int k = 0; k < kBvhMaxLevels; ++k)
    {
        shape.offset[k] = lay.offset[k];
        shape.count[k]  = k <= lay.top ? lay.count[k] : 0;
    }
    shape.top = lay.top;
    uint32_t* faults = faultWord();
    if (!faults) return HIPCSTONE_HIP_ERROR;
    HaloArgs<K, T> A{prefixes, childOffsets, parents, nodeCenters, nodeSizes, searchCenters, searchSizes,
                     leaves,   *box,         firstNode, lastNode,  numNodes,  flags,         faults};

    haloCountersInitKernel<<<1, 64, 0, stream(s)>>>(ctr, shape);
    HC_CHECK_LAUNCH("haloCountersInitKernel");
    const int32_t numLocal = lastNode - firstNode;
    haloLeafBoundsKernel<K, T><<<ceilDiv(size_t(numLocal), 256), 256, 0, stream(s)>>>(A, bvh, ctr);
    HC_CHECK_LAUNCH("haloLeafBoundsKernel");
    for (int k = 0; k < lay.top; ++k)
    {
        haloBoundsReduceKernel<T><<<ceilDiv(size_t(lay.count[k + 1]), 256), 256, 0, stream(s)>>>(
            bvh + size_t(lay.offset[k]) * 6, lay.count[k], bvh + size_t(lay.offset[k + 1]) * 6,
            lay.count[k + 1]);
        HC_CHECK_LAUNCH("haloBoundsReduceKernel");
    }
    haloRootKernel<K, T><<<1, 64, 0, stream(s)>>>(A, front, ctr);
    HC_CHECK_LAUNCH("haloRootKernel");
    for (int d = 0; d < int(maxLevel<K>()); ++d)
    {
        haloLevelKernel<K, T><<<kHaloBlocks, kHaloThreads, 0, stream(s)>>>(A, bvh, front, uint32_t(numNodes), ctr,
                                                                          d);
        HC_CHECK_LAUNCH("haloLevelKernel");
    }
    if (owned) HC_CHECK(hipFreeAsync(owned, stream(s)));
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

#define HIPCSTONE_DEF_BBOX(RS, T)                                                                                      \
    hipcstone_status hipcstone_compute_bounding_box_##RS(hipcstone_stream s, const T* x, const T* y, const T* z,      \
                                                         const T* h, const uint32_t* layout, int32_t first,           \
                                                         int32_t last, T scale, T* centers, T* sizes)                  \
    {                                                                                                                  \
        return boundingBox<T>(s, x, y, z, h, layout, first, last, scale, centers, sizes);                             \
    }

HIPCSTONE_DEF_BBOX(f32, float)
HIPCSTONE_DEF_BBOX(f64, double)

size_t hipcstone_find_halos_workspace_size(int32_t num_nodes, int32_t num_local_leaves, int32_t real_bytes)
{
    return haloLayout(num_nodes, num_local_leaves, real_bytes == 4 ? 4 : 8).bytes;
}

#define HIPCSTONE_DEF_HALOS(KS, K, RS, T)                                                                              \
    hipcstone_status hipcstone_find_halos_##KS##_##RS(                                                                 \
        hipcstone_stream s, const K* prefixes, const int32_t* child_offsets, const int32_t* parents,                  \
        const T* node_centers, const T* node_sizes, const K* leaves, const T* search_centers, const T* search_sizes,   \
        const hipcstone_box_##RS* box, int32_t first_node, int32_t last_node, int32_t num_nodes, uint8_t* flags,       \
        void* workspace, size_t workspace_bytes)                                                                       \
    {                                                                                                                  \
        return findHalos<K, T>(s, prefixes, child_offsets, parents, node_centers, node_sizes, leaves, search_centers, \
                               search_sizes, box, first_node, last_node, num_nodes, flags, workspace,                  \
                               workspace_bytes);                                                                       \
    }

HIPCSTONE_DEF_HALOS(u32, uint32_t, f32, float)
HIPCSTONE_DEF_HALOS(u32, uint32_t, f64, double)
HIPCSTONE_DEF_HALOS(u64, uint64_t, f32, float)
HIPCSTONE_DEF_HALOS(u64, uint64_t, f64, double)

} // extern "C"
