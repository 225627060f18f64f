// a10-a12: internal-tree linking, geometric node centers, leaf layout.
//
// Reference semantics: tree/octree.hpp:55-65 (binaryKeyWeight), :77-100 (createUnsortedLayout), :114-148 (linkTree),
// :151-160 (getLevelRange), :167-196 (buildOctree); focus/source_center.hpp:145-158 (nodeFpCenters) with
// sfc/box.hpp:318-335 (centerAndSize). The node order is level-major, SFC order within a level (sorted
// Warren-Salmon prefixes); every output array is uniquely determined by the leaves.
#include "common.hpp"
#include "runtime.hpp"
#include "scan.hpp"

extern "C" size_t hipcstone_sort_by_key_workspace_size_u32(size_t n);

namespace hipcstone
{

// createUnsortedLayout (octree.hpp:77-100): leaf tid is node tid + numInternal of the unsorted layout and, when it
// and the next leaf share a prefix of whole octal digits, creates internal node (tid + binaryKeyWeight) / 7. Besides
// the unsorted prefixes, every leaf emits its nodes' levels in creation order: entry 2 tid the internal node's (255
// when tid creates none), entry 2 tid + 1 its own. Within one level, creation order is key order (keys ascend with
// tid, and a level holds no node twice), so a stable sort of these entries by level alone gives the level-major,
// key-ordered node order that sorting the prefixes gives (a level's prefixes are exactly the ones with their
// placeholder bit at 3 * level).
template<class K>
__global__ __launch_bounds__(256) void nodeLevelsKernel(const K* __restrict__ leaves, int32_t numInternal,
                                                        int32_t numLeaves, K* __restrict__ unsortedPrefixes,
                                                        uint32_t* __restrict__ level, uint32_t* __restrict__ node)
{
    int32_t tid = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (tid >= numLeaves) return;
    K key                               = leaves[tid];
    K next                              = leaves[tid + 1];
    unsigned leafLevel                  = treeLevel(K(next - key));
    unsortedPrefixes[tid + numInternal] = encodePlaceholderBit(key, int(3 * leafLevel));
    level[2 * tid + 1]                  = leafLevel;
    node[2 * tid + 1]                   = uint32_t(tid + numInternal);
    unsigned prefixLength               = unsigned(commonPrefix(key, next));
    if (prefixLength % 3 == 0 && tid < numLeaves - 1)
    {
        int32_t octIndex           = (tid + binaryKeyWeight(key, prefixLength / 3)) / 7;
        unsortedPrefixes[octIndex] = encodePlaceholderBit(key, int(prefixLength));
        level[2 * tid]             = prefixLength / 3;
        node[2 * tid]              = uint32_t(octIndex);
    }
    else
    {
        level[2 * tid] = 255u;
        node[2 * tid]  = ~0u;
    }
}

// order = the first numNodes sorted entries: prefixes[i] = unsorted prefix of node order[i];
// leafToInternal[order[i]] = i; internalToLeaf[i] = order[i] - numInternal (octree.hpp:176-185)
template<class K>
__global__ void invertOrderKernel(const uint32_t* __restrict__ order, int32_t numNodes, int32_t numInternal,
                                  const K* __restrict__ unsortedPrefixes, K* __restrict__ prefixes,
                                  int32_t* __restrict__ internalToLeaf, int32_t* __restrict__ leafToInternal)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    int32_t o          = int32_t(order[i]);
    prefixes[i]        = unsortedPrefixes[o];
    leafToInternal[o]  = i;
    internalToLeaf[i]  = o - numInternal;
}

template<class K>
__device__ inline int32_t lowerBoundIdx(const K* a, int32_t lo, int32_t hi, K v)
{
    int32_t count = hi - lo;
    while (count > 0)
    {
        int32_t step = count / 2;
        int32_t it   = lo + step;
        if (a[it] < v)
        {
            lo = it + 1;
            count -= step + 1;
        }
        else
            count = step;
    }
    return lo;
}

// getLevelRangeCpu (octree.hpp:151-160): first node index of every level, levelRange[L+1] = numNodes
template<class K>
__global__ void levelRangeKernel(const K* __restrict__ prefixes, int32_t numNodes, int32_t* __restrict__ levelRange)
{
    int level = int(threadIdx.x);
    constexpr int L = int(maxLevel<K>());
    if (level <= L) levelRange[level] = lowerBoundIdx(prefixes, 0, numNodes, encodePlaceholderBit(K(0), 3 * level));
    if (level == L + 1) levelRange[level] = numNodes;
}

// linkTreeCpu (octree.hpp:114-148), childOffsets pre-zeroed
template<class K>
__global__ __launch_bounds__(256) void linkKernel(const K* __restrict__ prefixes, int32_t numInternal,
                                                  const int32_t* __restrict__ leafToInternal,
                                                  const int32_t* __restrict__ levelRange,
                                                  int32_t* __restrict__ childOffsets, int32_t* __restrict__ parents)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= numInternal) return;
    int32_t idxA          = leafToInternal[i];
    K prefix              = prefixes[idxA];
    K nodeKey             = decodePlaceholderBit(prefix);
    unsigned prefixLength = decodePrefixLength(prefix);
    unsigned level        = prefixLength / 3;
    K childPrefix         = encodePlaceholderBit(nodeKey, int(prefixLength) + 3);
    int32_t lo            = levelRange[level + 1];
    int32_t hi            = levelRange[level + 2];
    int32_t childIdx      = lowerBoundIdx(prefixes, lo, hi, childPrefix);
    if (childIdx != hi && childPrefix == prefixes[childIdx])
    {
        childOffsets[idxA]          = childIdx;
        parents[(childIdx - 1) / 8] = idxA;
    }
}

__global__ void fillI32Kernel(int32_t* p, size_t n, int32_t v)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// nodeFpCenters + centerAndSize; node keys decoded as Hilbert keys like the reference (sfcKey -> SfcKind)
template<class K, class T>
__global__ __launch_bounds__(256) void geoCentersKernel(const K* __restrict__ prefixes, int32_t numNodes,
                                                        T* __restrict__ centers, T* __restrict__ sizes,
                                                        typename BoxOf<T>::type box)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    K prefix       = prefixes[i];
    K startKey     = decodePlaceholderBit(prefix);
    unsigned level = decodePrefixLength(prefix) / 3;
    int ib[6];
    hilbertIBox(startKey, level, ib);
    constexpr T uL = T(1.) / T(1u << maxLevel<K>());
    T hx           = T(0.5) * uL * box.len[0];
    T hy           = T(0.5) * uL * box.len[1];
    T hz           = T(0.5) * uL * box.len[2];
    size_t o       = size_t(i) * 3;
    centers[o + 0] = box.lim[0] + T(ib[1] + ib[0]) * hx;
    centers[o + 1] = box.lim[2] + T(ib[3] + ib[2]) * hy;
    centers[o + 2] = box.lim[4] + T(ib[5] + ib[4]) * hz;
    sizes[o + 0]   = T(ib[1] - ib[0]) * hx;
    sizes[o + 1]   = T(ib[3] - ib[2]) * hy;
    sizes[o + 2]   = T(ib[5] - ib[4]) * hz;
}

extern "C" hipcstone_status hipcstone_sort_by_key_u32(hipcstone_stream, uint32_t*, uint32_t*, size_t, uint32_t*,
                                                      uint32_t*, void*, size_t, int);

static inline size_t a256(size_t v) { return (v + 255) & ~size_t(255); }

// workspace: [unsorted prefixes K (N)][levels u32 (2L)][nodes u32 (2L)][key buffer u32 (2L)][value buffer u32 (2L)]
// [sort workspace of 2L u32 keys]
template<class K>
static size_t buildWorkspace(int32_t numLeaves)
{
    size_t N = size_t(numLeaves) + size_t(numLeaves - 1) / 7;
    size_t E = 2 * size_t(numLeaves);
    return a256(N * sizeof(K)) + 4 * a256(E * 4) + hipcstone_sort_by_key_workspace_size_u32(E);
}

template<class K>
static hipcstone_status buildOctree(hipcstone_stream s, const K* leaves, int32_t numLeaves, K* prefixes,
                                    int32_t* childOffsets, int32_t* parents, int32_t* levelRange,
                                    int32_t* internalToLeaf, int32_t* leafToInternal, void* ws, size_t wsBytes,
                                    int32_t* levelRangeHost)
{
    if (numLeaves <= 0) return invalid("buildOctree: numLeaves must be positive");
    if (!leaves || !prefixes || !childOffsets || !parents || !levelRange || !internalToLeaf || !leafToInternal)
        return invalid("buildOctree: null pointer");
    if (!ws || wsBytes < buildWorkspace<K>(numLeaves)) return invalid("buildOctree: workspace too small");
    hipStream_t st      = stream(s);
    int32_t numInternal = (numLeaves - 1) / 7;
    int32_t numNodes    = numInternal + numLeaves;
    size_t N            = size_t(numNodes);
    size_t E            = 2 * size_t(numLeaves);
    char* w             = static_cast<char*>(ws);
    K* unsortedPrefixes = reinterpret_cast<K*>(w);
    w += a256(N * sizeof(K));
    uint32_t* level = reinterpret_cast<uint32_t*>(w);
    w += a256(E * 4);
    uint32_t* order = reinterpret_cast<uint32_t*>(w); // the node of each entry, then (sorted) the node order
    w += a256(E * 4);
    uint32_t* keyBuf = reinterpret_cast<uint32_t*>(w);
    w += a256(E * 4);
    uint32_t* valueBuf = reinterpret_cast<uint32_t*>(w);
    w += a256(E * 4);
    const size_t sortWsBytes = wsBytes - size_t(w - static_cast<char*>(ws));

    nodeLevelsKernel<K><<<ceilDiv(size_t(numLeaves), 256), 256, 0, st>>>(leaves, numInternal, numLeaves,
                                                                          unsortedPrefixes, level, order);
    HC_CHECK_LAUNCH("nodeLevelsKernel");
    // one 8-bit pass: levels are at most 21, 255 marks the entries without a node (they sort last)
    if (auto e = hipcstone_sort_by_key_u32(st, level, order, E, keyBuf, valueBuf, w, sortWsBytes, 8)) return e;
    invertOrderKernel<K><<<ceilDiv(N, 256), 256, 0, st>>>(order, numNodes, numInternal, unsortedPrefixes, prefixes,
                                                          internalToLeaf, leafToInternal);
    HC_CHECK_LAUNCH("invertOrderKernel");
    levelRangeKernel<K><<<1, 64, 0, st>>>(prefixes, numNodes, levelRange);
    HC_CHECK_LAUNCH("levelRangeKernel");
    fillI32Kernel<<<ceilDiv(N, 256), 256, 0, st>>>(childOffsets, N, 0);
    HC_CHECK_LAUNCH("fillI32Kernel");
    if (numInternal > 0)
    {
        linkKernel<K><<<ceilDiv(size_t(numInternal), 256), 256, 0, st>>>(prefixes, numInternal, leafToInternal,
                                                                         levelRange, childOffsets, parents);
        HC_CHECK_LAUNCH("linkKernel");
    }
    if (levelRangeHost)
    {
        HC_CHECK(hipMemcpyAsync(levelRangeHost, levelRange, sizeof(int32_t) * (maxLevel<K>() + 2),
                                hipMemcpyDeviceToHost, st));
        HC_CHECK(hipStreamSynchronize(st));
    }
    return HIPCSTONE_OK;
}

template<class K, class T>
static hipcstone_status geoCenters(hipcstone_stream s, const K* prefixes, int32_t numNodes, T* centers, T* sizes,
                                   const typename BoxOf<T>::type* box)
{
    if (numNodes <= 0) return HIPCSTONE_OK;
    if (!prefixes || !centers || !sizes || !box) return invalid("computeGeoCenters: null pointer");
    geoCentersKernel<K, T><<<ceilDiv(size_t(numNodes), 256), 256, 0, stream(s)>>>(prefixes, numNodes, centers, sizes,
                                                                                   *box);
    HC_CHECK_LAUNCH("geoCentersKernel");
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

size_t hipcstone_build_octree_workspace_size_u32(int32_t num_leaves) { return buildWorkspace<uint32_t>(num_leaves); }
size_t hipcstone_build_octree_workspace_size_u64(int32_t num_leaves) { return buildWorkspace<uint64_t>(num_leaves); }

hipcstone_status hipcstone_build_octree_u32(hipcstone_stream s, const uint32_t* leaves, int32_t num_leaves,
                                            uint32_t* prefixes, int32_t* child_offsets, int32_t* parents,
                                            int32_t* level_range, int32_t* internal_to_leaf,
                                            int32_t* leaf_to_internal, void* workspace, size_t workspace_bytes,
                                            int32_t* level_range_host)
{
    return buildOctree<uint32_t>(s, leaves, num_leaves, prefixes, child_offsets, parents, level_range,
                                 internal_to_leaf, leaf_to_internal, workspace, workspace_bytes, level_range_host);
}

hipcstone_status hipcstone_build_octree_u64(hipcstone_stream s, const uint64_t* leaves, int32_t num_leaves,
                                            uint64_t* prefixes, int32_t* child_offsets, int32_t* parents,
                                            int32_t* level_range, int32_t* internal_to_leaf,
                                            int32_t* leaf_to_internal, void* workspace, size_t workspace_bytes,
                                            int32_t* level_range_host)
{
    return buildOctree<uint64_t>(s, leaves, num_leaves, prefixes, child_offsets, parents, level_range,
                                 internal_to_leaf, leaf_to_internal, workspace, workspace_bytes, level_range_host);
}

#define HIPCSTONE_DEF_CENTERS(KS, K, RS, T)                                                                            \
    hipcstone_status hipcstone_compute_geo_centers_##KS##_##RS(hipcstone_stream s, const K* prefixes,                 \
                                                               int32_t num_nodes, T* centers, T* sizes,                \
                                                               const hipcstone_box_##RS* box)                          \
    {                                                                                                                  \
        return geoCenters<K, T>(s, prefixes, num_nodes, centers, sizes, box);                                         \
    }

HIPCSTONE_DEF_CENTERS(u32, uint32_t, f32, float)
HIPCSTONE_DEF_CENTERS(u32, uint32_t, f64, double)
HIPCSTONE_DEF_CENTERS(u64, uint64_t, f32, float)
HIPCSTONE_DEF_CENTERS(u64, uint64_t, f64, double)

size_t hipcstone_scan_workspace_size(size_t n) { return scan::workspaceBytes(n); }

hipcstone_status hipcstone_leaf_layout(hipcstone_stream s, const uint32_t* counts, int32_t num_leaves,
                                       uint32_t* layout, void* workspace, size_t workspace_bytes)
{
    if (num_leaves <= 0) return invalid("leafLayout: num_leaves must be positive");
    if (!counts || !layout) return invalid("leafLayout: null pointer");
    return scan::exclusiveScan<uint32_t>(stream(s), counts, layout, size_t(num_leaves), 0u, layout + num_leaves,
                                         workspace, workspace_bytes);
}

} // extern "C"
