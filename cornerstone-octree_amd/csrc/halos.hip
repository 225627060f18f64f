// a13/a14: per-leaf search boxes and halo discovery.
//
// Reference semantics: focus/source_center.hpp:28-43 (computeBoundingBox) and traversal/collisions.hpp:25-94
// (findCollisions, findHalos) with traversal/boxoverlap.hpp:127-168 (containedIn) and :280-291 (overlap).
// GPU interface: focus/source_center_gpu.h:40-51, traversal/collisions_gpu.h:46-59.
#include <algorithm>

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
// when it returns true and descends into the node's children only then. So leaf l flags node n iff overlaps() holds
// for l at n and at every ancestor of n; leaves whose search box lies inside the local key range walk nothing.
//
// GPU design. The 64 consecutive local leaves of a group share most of their walks, so a wave walks the union of a
// group's walks: expanding a node tests its 8 children at once (their boxes, prefixes and child offsets are contiguous
// and wave-uniform) for the lanes that passed the node and keeps the per-child ballot of passing lanes. A child with a
// non-empty ballot is flagged (one byte store per node and group instead of one per leaf) and, if internal, becomes a
// frame (group, first child, ballot). Each lane is tested exactly where its own reference walk tests, so the flag set
// is the reference's.
//
// Union walks differ in length by orders of magnitude (a group next to a dense cluster expands tens of thousands of
// frames, the median group a few hundred), so one wave per group leaves the GPU waiting on a few waves. The frames are
// instead expanded level by level: level d reads the frames of depth-d nodes from a queue and appends the frames of
// their children to the next queue, with every wave of a fixed grid taking frames round-robin. Waves stage their
// output frames in LDS and reserve queue space with one atomic per batch. A frame that finds its queue full is walked
// to the bottom by the wave that made it (depth first, with an explicit stack in LDS), so any queue capacity gives the
// same flags; the capacity only decides how much of the work is balanced.
constexpr int kHaloThreads     = 256;
constexpr int kHaloWaves       = kHaloThreads / kWave;
constexpr int kHaloMaxDepth    = 22;
constexpr int kHaloStage       = 64;   // staged output frames per wave
constexpr unsigned kHaloBlocks = 2048; // grid of the level kernels (8 waves per SIMD)
constexpr size_t kHaloCounters = 32;   // uint64 queue counters at the start of the workspace, one per level

struct HaloFrame
{
    uint64_t mask; // lanes of the group that passed the parent node
    int32_t group; // leaf group: local leaves firstNode + 64 * group + lane
    int32_t co;    // first of the 8 children to test
};

struct HaloShared
{
    HaloFrame stage[kHaloWaves][kHaloStage];
    uint64_t mask[kHaloWaves][kHaloMaxDepth][8];
    int32_t cco[kHaloWaves][kHaloMaxDepth][8]; // first child of each of the frame's 8 nodes
    int32_t co[kHaloWaves][kHaloMaxDepth];
    int32_t next[kHaloWaves][kHaloMaxDepth];
};

template<class K, class T>
struct HaloTree
{
    const K* prefixes;
    const int32_t* childOffsets;
    const T* nodeCenters;
    const T* nodeSizes;
    const T* searchCenters;
    const T* searchSizes;
    typename BoxOf<T>::type box;
    int32_t firstNode, lastNode, numNodes;
    uint8_t* flags;
};

// Per-lane state of one group: the lane's search box and the local key range.
template<class K, class T>
struct HaloLane
{
    T tC[3], tS[3];
    K lowestKey, highestKey;
    bool pbc[3];

    __device__ void init(const HaloTree<K, T>& t, K lo, K hi)
    {
        lowestKey  = lo;
        highestKey = hi;
        for (int d = 0; d < 3; ++d)
            pbc[d] = t.box.bnd[d] == HIPCSTONE_PERIODIC;
    }

    // the lane's search box for a group; lanes outside the mask keep stale values that are never used
    __device__ void load(const HaloTree<K, T>& t, int32_t group, uint64_t mask, int lane)
    {
        if ((mask >> lane) & 1u)
        {
            size_t leaf = size_t(t.firstNode) + size_t(group) * kWave + size_t(lane);
            for (int d = 0; d < 3; ++d)
            {
                tC[d] = t.searchCenters[leaf * 3 + d];
                tS[d] = t.searchSizes[leaf * 3 + d];
            }
        }
    }

    // containedIn(nodeStart, nodeEnd, lowestKey, highestKey) of the node with this prefix (boxoverlap.hpp:163-168)
    __device__ bool excluded(K pfx) const
    {
        unsigned pl = decodePrefixLength(pfx);
        K ret       = pfx ^ (K(1) << pl);
        int ns      = 3 * int(maxLevel<K>()) - int(pl);
        K nk1       = ret << ns;
        K nk2       = nk1 + (K(1) << ns);
        return !(nk1 < lowestKey || nk2 > highestKey);
    }

    // overlap(node box, search box) with the reference's operation order (boxoverlap.hpp:280-291)
    __device__ bool overlapBox(const T* c, const T* s, const typename BoxOf<T>::type& box) const
    {
        bool ov = true;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            T dx = tC[d] - c[d];
            if (pbc[d]) dx = dx - box.len[d] * rint(dx * box.ilen[d]);
            dx = fabs(dx);
            dx -= s[d];
            dx -= tS[d];
            ov = ov && dx < T(0);
        }
        return ov;
    }
};

// Test the 8 children at co for the lanes in pm: mk[k] = ballot of lanes for which child k overlaps, cco[k] = its
// first child. Lanes 0-7 flag the children some lane passes.
template<class K, class T>
__device__ inline void testChildren(const HaloTree<K, T>& t, const HaloLane<K, T>& L, int32_t co, uint64_t pm,
                                    int lane, uint64_t mk[8], int32_t cco[8])
{
    const bool eligible = (pm >> lane) & 1u;
    const T* c          = t.nodeCenters + size_t(co) * 3;
    const T* s          = t.nodeSizes + size_t(co) * 3;
    T cv[24], sv[24];
#pragma unroll
    for (int k = 0; k < 24; ++k)
    {
        cv[k] = c[k];
        sv[k] = s[k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        bool ex = L.excluded(t.prefixes[co + k]);
        cco[k]  = t.childOffsets[co + k];
        mk[k]   = ballot(eligible && !ex && L.overlapBox(cv + 3 * k, sv + 3 * k, t.box));
    }
    if (lane < 8)
    {
        uint64_t mine = mk[0];
#pragma unroll
        for (int k = 1; k < 8; ++k)
            mine = lane == k ? mk[k] : mine;
        if (mine != 0) t.flags[co + lane] = 1;
    }
}

// child k of the frame at co is an internal node whose 8 children lie inside the tree (malformed trees stop here)
__device__ inline bool descendable(int32_t co, int k, int32_t cco, int32_t numNodes)
{
    return cco > co + k && cco + 8 <= numNodes;
}

// Depth-first walk of everything below one frame, for frames that found no room in the level queue.
template<class K, class T>
__device__ void walkFrame(const HaloTree<K, T>& t, HaloLane<K, T>& L, HaloShared& sh, int w, int lane,
                          const HaloFrame f)
{
    auto push = [&](int depth, int32_t co, uint64_t pm) {
        uint64_t mk[8];
        int32_t cco[8];
        testChildren(t, L, co, pm, lane, mk, cco);
        if (lane == 0)
        {
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                sh.mask[w][depth][k] = descendable(co, k, cco[k], t.numNodes) ? mk[k] : 0;
                sh.cco[w][depth][k]  = cco[k];
            }
            sh.co[w][depth]   = co;
            sh.next[w][depth] = 0;
        }
        waveSync();
    };

    L.load(t, f.group, f.mask, lane);
    int depth = 0;
    push(0, f.co, f.mask);
    int64_t stepsLeft = 2 * int64_t(t.numNodes) + 8;
    while (depth >= 0 && stepsLeft-- > 0)
    {
        int k      = uniform(sh.next[w][depth]);
        uint64_t m = 0;
        while (k < 8)
        {
            m = uniform64(sh.mask[w][depth][k]);
            if (m != 0) break;
            ++k;
        }
        if (k >= 8)
        {
            --depth;
            continue;
        }
        int32_t cco = uniform(sh.cco[w][depth][k]);
        if (lane == 0) sh.next[w][depth] = k + 1;
        waveSync();
        if (depth + 1 >= kHaloMaxDepth) break; // deeper than any key allows: malformed tree
        ++depth;
        push(depth, cco, m);
    }
}

// Root test of every group (collisions.hpp:33-34 via singleTraversal, traversal.hpp:35-45; the containedIn skip of
// findHalos :88-89): one wave per group writes the group's frame for the root's children, with an empty ballot when
// the group walks nothing. Block 0 also resets the level counters.
template<class K, class T>
__global__ __launch_bounds__(kHaloThreads) void haloRootKernel(HaloTree<K, T> t, const K* __restrict__ leaves,
                                                               HaloFrame* __restrict__ out,
                                                               unsigned long long* __restrict__ counters,
                                                               int32_t numGroups)
{
    if (blockIdx.x == 0 && threadIdx.x < kHaloCounters)
        counters[threadIdx.x] = threadIdx.x == 1 ? (unsigned long long)numGroups : 0ull;
    const int lane    = laneId();
    const int32_t g   = int32_t((size_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave);
    if (g >= numGroups) return;
    const int32_t leafIdx = t.firstNode + g * kWave + lane;
    HaloLane<K, T> L;
    L.init(t, leaves[t.firstNode], leaves[t.lastNode]);
    bool active = leafIdx < t.lastNode;
    if (active)
    {
        for (int d = 0; d < 3; ++d)
        {
            L.tC[d] = t.searchCenters[size_t(leafIdx) * 3 + d];
            L.tS[d] = t.searchSizes[size_t(leafIdx) * 3 + d];
        }
        active = !boxContainedInRange<K, T>(L.lowestKey, L.highestKey, L.tC, L.tS, t.box);
    }
    uint64_t rootMask = ballot(active && !L.excluded(t.prefixes[0]) && L.overlapBox(t.nodeCenters, t.nodeSizes, t.box));
    int32_t rootChild = t.childOffsets[0];
    if (rootMask != 0 && lane == 0) t.flags[0] = 1;
    if (!descendable(0, 0, rootChild, t.numNodes)) rootMask = 0;
    if (lane == 0) out[g] = HaloFrame{rootMask, g, rootChild};
}

// One level of the union walks: expands the frames in[0, min(*inCount, cap)) and appends the children's frames to
// out, walking depth first the frames that find out full.
template<class K, class T>
__global__ __launch_bounds__(kHaloThreads) void haloLevelKernel(HaloTree<K, T> t, const K* __restrict__ leaves,
                                                                const HaloFrame* __restrict__ in,
                                                                const unsigned long long* __restrict__ inCount,
                                                                HaloFrame* __restrict__ out,
                                                                unsigned long long* __restrict__ outCount,
                                                                unsigned long long cap)
{
    __shared__ HaloShared sh;
    const int lane = laneId();
    const int w    = threadIdx.x / kWave;
    const uint64_t n     = uniform64(*inCount < cap ? *inCount : cap);
    const uint64_t waves = uint64_t(gridDim.x) * kHaloWaves;
    const uint64_t gw    = uint64_t(xcdContiguousBlock(blockIdx.x, gridDim.x)) * kHaloWaves + uint64_t(w);
    if (gw >= n) return;
    HaloLane<K, T> L;
    L.init(t, leaves[t.firstNode], leaves[t.lastNode]);
    int staged = 0;

    auto flush = [&]() {
        waveSync();
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(outCount, (unsigned long long)staged);
        base = uniform64(base);
        if (lane < staged && base + lane < cap) out[base + lane] = sh.stage[w][lane];
        for (uint64_t j = base < cap ? cap - base : 0; j < uint64_t(staged); ++j)
            walkFrame(t, L, sh, w, lane, sh.stage[w][j]);
        staged = 0;
        waveSync();
    };

    for (uint64_t it = gw; it < n; it += waves)
    {
        const HaloFrame f = in[it];
        const uint64_t pm = uniform64(f.mask);
        if (pm == 0) continue;
        const int32_t group = uniform(f.group);
        const int32_t co    = uniform(f.co);
        L.load(t, group, pm, lane);
        uint64_t mk[8];
        int32_t cco[8];
        testChildren(t, L, co, pm, lane, mk, cco);
        uint64_t mine = mk[0];
        int32_t myCo  = cco[0];
#pragma unroll
        for (int k = 1; k < 8; ++k)
        {
            mine = lane == k ? mk[k] : mine;
            myCo = lane == k ? cco[k] : myCo;
        }
        const bool emit    = lane < 8 && mine != 0 && descendable(co, lane, myCo, t.numNodes);
        const uint64_t eb  = ballot(emit);
        const int rank     = __builtin_popcountll(eb & ((uint64_t(1) << lane) - 1));
        if (emit) sh.stage[w][staged + rank] = HaloFrame{mine, group, myCo};
        staged += __builtin_popcountll(eb);
        if (staged > kHaloStage - 8) flush();
    }
    if (staged > 0) flush();
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

// Workspace of findHalos: the level counters, then two frame queues of equal capacity.
inline unsigned long long haloQueueCapacity(size_t workspaceBytes)
{
    size_t head = kHaloCounters * sizeof(unsigned long long);
    return workspaceBytes <= head ? 0 : (workspaceBytes - head) / (2 * sizeof(HaloFrame));
}

inline size_t haloWorkspaceBytes(size_t queueCapacity)
{
    return kHaloCounters * sizeof(unsigned long long) + 2 * queueCapacity * sizeof(HaloFrame);
}

// Recommended size: queues of two frames per tree node (frames beyond the capacity are walked depth first).
inline size_t findHalosWorkspaceSize(int32_t numNodes)
{
    size_t cap = std::min(std::max(size_t(2) * size_t(std::max(numNodes, 0)), size_t(1) << 14), size_t(1) << 25);
    return haloWorkspaceBytes(cap);
}

// Resident blocks of the level kernel over the whole GPU, a multiple of the XCD count.
template<class K, class T>
static unsigned haloLevelGrid()
{
    static unsigned cached[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cached[dev]) return cached[dev];
    int perCu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, haloLevelKernel<K, T>, kHaloThreads, 0) != hipSuccess ||
        perCu <= 0)
        perCu = 2;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
    unsigned g  = unsigned(perCu) * unsigned(cus);
    g           = std::min(std::max(g, kXcds), kHaloBlocks * 4);
    cached[dev] = (g + kXcds - 1) / kXcds * kXcds;
    return cached[dev];
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
    const size_t groups = (size_t(lastNode - firstNode) + kWave - 1) / kWave;
    void* owned         = nullptr;
    if (!workspace)
    {
        workspaceBytes = std::max(findHalosWorkspaceSize(numNodes), haloWorkspaceBytes(std::max(groups, size_t(64))));
        HC_CHECK(hipMallocAsync(&owned, workspaceBytes, stream(s)));
        workspace = owned;
    }
    const unsigned long long cap = haloQueueCapacity(workspaceBytes);
    if (cap < std::max(groups, size_t(kHaloStage)))
        return invalid("findHalos: workspace smaller than 256 + 32 * max(64, ceil((last - first) / 64)) bytes");

    auto* counters = static_cast<unsigned long long*>(workspace);
    auto* queueA   = reinterpret_cast<HaloFrame*>(counters + kHaloCounters);
    auto* queueB   = queueA + cap;
    HaloTree<K, T> t{prefixes, childOffsets, nodeCenters, nodeSizes, searchCenters, searchSizes, *box,
                     firstNode,  lastNode,    numNodes,   flags};
    haloRootKernel<K, T><<<ceilDiv(groups * kWave, kHaloThreads), kHaloThreads, 0, stream(s)>>>(
        t, leaves, queueA, counters, int32_t(groups));
    HC_CHECK_LAUNCH("haloRootKernel");
    const unsigned grid = haloLevelGrid<K, T>();
    for (int d = 1; d <= int(maxLevel<K>()); ++d)
    {
        HaloFrame* in  = (d & 1) ? queueA : queueB;
        HaloFrame* out = (d & 1) ? queueB : queueA;
        haloLevelKernel<K, T><<<grid, kHaloThreads, 0, stream(s)>>>(t, leaves, in, counters + d, out,
                                                                    counters + d + 1, cap);
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

size_t hipcstone_find_halos_workspace_size(int32_t num_nodes) { return findHalosWorkspaceSize(num_nodes); }

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
