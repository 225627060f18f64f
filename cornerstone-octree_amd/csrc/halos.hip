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

// findHalos: one lane per local leaf, each with its own stackless walk (collisions.hpp:25-48). Flags only ever
// receive 1, so concurrent writers agree. The walk is capped at 2 * numNodes + 8 steps.
template<class K, class T>
__global__ __launch_bounds__(128) void findHalosKernel(const K* __restrict__ prefixes,
                                                       const int32_t* __restrict__ childOffsets,
                                                       const int32_t* __restrict__ parents,
                                                       const T* __restrict__ nodeCenters,
                                                       const T* __restrict__ nodeSizes, const K* __restrict__ leaves,
                                                       const T* __restrict__ searchCenters,
                                                       const T* __restrict__ searchSizes,
                                                       typename BoxOf<T>::type box, int32_t firstNode,
                                                       int32_t lastNode, int32_t numNodes, uint8_t* flags)
{
    int32_t leafIdx = firstNode + int32_t(size_t(blockIdx.x) * blockDim.x + threadIdx.x);
    if (leafIdx >= lastNode) return;
    K lowestKey  = leaves[firstNode];
    K highestKey = leaves[lastNode];
    T tC[3] = {searchCenters[size_t(leafIdx) * 3], searchCenters[size_t(leafIdx) * 3 + 1],
               searchCenters[size_t(leafIdx) * 3 + 2]};
    T tS[3] = {searchSizes[size_t(leafIdx) * 3], searchSizes[size_t(leafIdx) * 3 + 1],
               searchSizes[size_t(leafIdx) * 3 + 2]};
    if (boxContainedInRange<K, T>(lowestKey, highestKey, tC, tS, box)) return;

    const bool pbc[3] = {box.bnd[0] == HIPCSTONE_PERIODIC, box.bnd[1] == HIPCSTONE_PERIODIC,
                         box.bnd[2] == HIPCSTONE_PERIODIC};
    auto overlaps = [&](int32_t idx) -> bool {
        K pfx        = prefixes[idx];
        unsigned pl  = decodePrefixLength(pfx);
        K ret        = pfx ^ (K(1) << pl);
        int ns       = 3 * int(maxLevel<K>()) - int(pl);
        K nk1        = ret << ns;
        K nk2        = nk1 + (K(1) << ns);
        bool contained = !(nk1 < lowestKey || nk2 > highestKey);
        if (contained) return false;
        bool ov = true;
#pragma unroll
        for (int d = 0; d < 3; ++d)
        {
            T dx = tC[d] - nodeCenters[size_t(idx) * 3 + d];
            if (pbc[d]) dx = dx - box.len[d] * rint(dx * box.ilen[d]);
            dx = fabs(dx);
            dx -= nodeSizes[size_t(idx) * 3 + d];
            dx -= tS[d];
            ov = ov && dx < T(0);
        }
        if (ov) flags[idx] = 1;
        return ov;
    };

    if (!overlaps(0)) return;
    if (childOffsets[0] == 0) return;
    int32_t node      = childOffsets[0];
    bool backtrack    = false;
    int64_t stepsLeft = 2 * int64_t(numNodes) + 8;
    while (node != 0 && stepsLeft-- > 0)
    {
        if (node < 0 || node >= numNodes) break;
        bool isLeaf  = childOffsets[node] == 0;
        bool descend = !backtrack && overlaps(node);
        int32_t sib  = (node - 1) % 8;
        if (!isLeaf && descend)
        {
            node      = childOffsets[node];
            backtrack = false;
        }
        else if (sib < 7)
        {
            node++;
            backtrack = false;
        }
        else
        {
            node      = parents[(node - 1) / 8];
            backtrack = true;
        }
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
    findHalosKernel<K, T><<<ceilDiv(size_t(lastNode - firstNode), 128), 128, 0, stream(s)>>>(
        prefixes, childOffsets, parents, nodeCenters, nodeSizes, leaves, searchCenters, searchSizes, *box, firstNode,
        lastNode, numNodes, flags);
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
