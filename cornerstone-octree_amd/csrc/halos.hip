// a13/a14: per-leaf search boxes and halo discovery.
//
// Reference semantics: focus/source_center.hpp:28-43 (computeBoundingBox) and traversal/collisions.hpp:25-94
// (findCollisions, findHalos) with traversal/boxoverlap.hpp:127-168 (containedIn) and :280-291 (overlap).
// GPU interface: focus/source_center_gpu.h:40-51, traversal/collisions_gpu.h:46-59.
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
// GPU design: one wave owns 64 consecutive local leaves, which are close on the SFC and share most of their walks.
// The wave walks the union depth-first with an explicit stack in LDS; expanding a node tests its 8 children at once
// (their boxes, prefixes and child offsets are contiguous and wave-uniform) for the lanes that passed the node, and
// keeps the per-child ballot of passing lanes. A child with a non-empty ballot is flagged (one byte store per node and
// wave instead of one per leaf) and, if internal, expanded with that ballot. Each lane is tested exactly where its own
// reference walk tests, so the flag set is the reference's. Every internal node is expanded at most once per wave,
// and the stack depth is bounded by the key's level count.
constexpr int kHaloThreads  = 256;
constexpr int kHaloWaves    = kHaloThreads / kWave;
constexpr int kHaloMaxDepth = 22;

struct HaloShared
{
    uint64_t mask[kHaloWaves][kHaloMaxDepth][8];
    int32_t cco[kHaloWaves][kHaloMaxDepth][8]; // first child of each of the frame's 8 nodes
    int32_t co[kHaloWaves][kHaloMaxDepth];
    int32_t next[kHaloWaves][kHaloMaxDepth];
};

template<class K, class T>
__global__ __launch_bounds__(kHaloThreads) void findHalosKernel(const K* __restrict__ prefixes,
                                                                const int32_t* __restrict__ childOffsets,
                                                                const T* __restrict__ nodeCenters,
                                                                const T* __restrict__ nodeSizes,
                                                                const K* __restrict__ leaves,
                                                                const T* __restrict__ searchCenters,
                                                                const T* __restrict__ searchSizes,
                                                                typename BoxOf<T>::type box, int32_t firstNode,
                                                                int32_t lastNode, int32_t numNodes, uint8_t* flags,
                                                                uint32_t numBlocksPadded)
{
    __shared__ HaloShared sh;
    const int lane       = laneId();
    const int w          = threadIdx.x / kWave;
    const uint32_t block = xcdContiguousBlock(blockIdx.x, numBlocksPadded);
    const int64_t base   = int64_t(firstNode) + int64_t(uniform(block * uint32_t(kHaloWaves) + uint32_t(w))) * kWave;
    if (base >= lastNode) return;
    const int32_t leafIdx = int32_t(base) + lane;
    const K lowestKey     = leaves[firstNode];
    const K highestKey    = leaves[lastNode];

    T tC[3] = {0, 0, 0}, tS[3] = {0, 0, 0};
    bool active = leafIdx < lastNode;
    if (active)
    {
        for (int d = 0; d < 3; ++d)
        {
            tC[d] = searchCenters[size_t(leafIdx) * 3 + d];
            tS[d] = searchSizes[size_t(leafIdx) * 3 + d];
        }
        active = !boxContainedInRange<K, T>(lowestKey, highestKey, tC, tS, box);
    }
    const bool pbc[3] = {box.bnd[0] == HIPCSTONE_PERIODIC, box.bnd[1] == HIPCSTONE_PERIODIC,
                         box.bnd[2] == HIPCSTONE_PERIODIC};

    // containedIn(nodeStart, nodeEnd, lowestKey, highestKey) of the node with this prefix (boxoverlap.hpp:163-168)
    auto excluded = [&](K pfx) -> bool {
        unsigned pl = decodePrefixLength(pfx);
        K ret       = pfx ^ (K(1) << pl);
        int ns      = 3 * int(maxLevel<K>()) - int(pl);
        K nk1       = ret << ns;
        K nk2       = nk1 + (K(1) << ns);
        return !(nk1 < lowestKey || nk2 > highestKey);
    };
    // overlap(node box, search box) with the reference's operation order (boxoverlap.hpp:280-291)
    auto overlapBox = [&](const T* c, const T* s) -> bool {
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
    };

    // test the 8 children of a node for the lanes in pm, flag those some lane passes, push the internal ones
    auto pushFrame = [&](int depth, int32_t co, uint64_t pm) {
        const bool eligible = (pm >> lane) & 1u;
        const T* c          = nodeCenters + size_t(co) * 3;
        const T* s          = nodeSizes + size_t(co) * 3;
        T cv[24], sv[24];
#pragma unroll
        for (int k = 0; k < 24; ++k)
        {
            cv[k] = c[k];
            sv[k] = s[k];
        }
        uint64_t mk[8];
        int32_t cco[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            bool ex = excluded(prefixes[co + k]);
            cco[k]  = childOffsets[co + k];
            mk[k]   = ballot(eligible && !ex && overlapBox(cv + 3 * k, sv + 3 * k));
        }
        if (lane < 8)
        {
            uint64_t mine = mk[0];
#pragma unroll
            for (int k = 1; k < 8; ++k)
                mine = lane == k ? mk[k] : mine;
            if (mine != 0) flags[co + lane] = 1;
        }
        if (lane == 0)
        {
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                sh.mask[w][depth][k] = cco[k] != 0 ? mk[k] : 0;
                sh.cco[w][depth][k]  = cco[k];
            }
            sh.co[w][depth]   = co;
            sh.next[w][depth] = 0;
        }
        waveSync();
    };

    // root (collisions.hpp:33-34 via singleTraversal, traversal.hpp:35-45)
    uint64_t rootMask = ballot(active && !excluded(prefixes[0]) && overlapBox(nodeCenters, nodeSizes));
    if (rootMask == 0) return;
    if (lane == 0) flags[0] = 1;
    int32_t rootChild = uniform(childOffsets[0]);
    if (rootChild == 0) return;
    if (rootChild < 1 || rootChild + 8 > numNodes) return; // malformed tree guard
    int depth = 0;
    pushFrame(0, rootChild, rootMask);
    int64_t stepsLeft = 2 * int64_t(numNodes) + 8;
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
        int32_t node = uniform(sh.co[w][depth]) + k;
        int32_t cco  = uniform(sh.cco[w][depth][k]);
        if (lane == 0) sh.next[w][depth] = k + 1;
        waveSync();
        if (node <= 0 || node >= numNodes) break; // malformed tree guard
        if (depth + 1 >= kHaloMaxDepth || cco <= node || cco + 8 > numNodes) break;
        ++depth;
        pushFrame(depth, cco, m);
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

template<class K, class T>
static hipcstone_status findHalos(hipcstone_stream s, const K* prefixes, const int32_t* childOffsets,
                                  const int32_t* parents, const T* nodeCenters, const T* nodeSizes, const K* leaves,
                                  const T* searchCenters, const T* searchSizes, const typename BoxOf<T>::type* box,
                                  int32_t firstNode, int32_t lastNode, int32_t numNodes, uint8_t* flags)
{
    if (lastNode <= firstNode) return HIPCSTONE_OK;
    if (!prefixes || !childOffsets || !parents || !nodeCenters || !nodeSizes || !leaves || !searchCenters ||
        !searchSizes || !box || !flags)
        return invalid("findHalos: null pointer");
    if (numNodes <= 0) return invalid("findHalos: numNodes must be positive");
    size_t groups   = (size_t(lastNode - firstNode) + kWave - 1) / kWave;
    unsigned blocks = ceilDiv(groups * kWave, kHaloThreads);
    blocks          = (blocks + kXcds - 1) / kXcds * kXcds; // whole rounds over the XCDs; extra groups exit
    findHalosKernel<K, T><<<blocks, kHaloThreads, 0, stream(s)>>>(prefixes, childOffsets, nodeCenters, nodeSizes,
                                                                  leaves, searchCenters, searchSizes, *box, firstNode,
                                                                  lastNode, numNodes, flags, blocks);
    HC_CHECK_LAUNCH("findHalosKernel");
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

#define HIPCSTONE_DEF_HALOS(KS, K, RS, T)                                                                              \
    hipcstone_status hipcstone_find_halos_##KS##_##RS(                                                                 \
        hipcstone_stream s, const K* prefixes, const int32_t* child_offsets, const int32_t* parents,                  \
        const T* node_centers, const T* node_sizes, const K* leaves, const T* search_centers, const T* search_sizes,   \
        const hipcstone_box_##RS* box, int32_t first_node, int32_t last_node, int32_t num_nodes, uint8_t* flags)       \
    {                                                                                                                  \
        return findHalos<K, T>(s, prefixes, child_offsets, parents, node_centers, node_sizes, leaves, search_centers, \
                               search_sizes, box, first_node, last_node, num_nodes, flags);                            \
    }

HIPCSTONE_DEF_HALOS(u32, uint32_t, f32, float)
HIPCSTONE_DEF_HALOS(u32, uint32_t, f64, double)
HIPCSTONE_DEF_HALOS(u64, uint64_t, f32, float)
HIPCSTONE_DEF_HALOS(u64, uint64_t, f64, double)

} // extern "C"
