// The reference-side drop-in of hipcstone (definitions). This translation unit replaces, in a cornerstone build, the
// .cu files that define the GPU functions of the path: sfc_gpu.cu, csarray_gpu.cu (counts, ops, rebalance),
// octree_gpu.cu (buildOctreeGpu), source_center_gpu.cu (bounding boxes, geometric centers, source centers, MAC
// spheres), collisions_gpu.cu (findHalosGpu), gather_halos_gpu.cu and the sort of primitives_gpu.cu. It defines the
// templates those *_gpu.h headers declare by calling libhipcstone's C ABI and instantiates them for the reference's
// own type sets (the .cu lines cited at each). It is ordinary host C++: compile it with the reference's headers and
// hipcstone.h on the include path and __HIP_PLATFORM_AMD__ defined (so execution::Gpu converts to a hipStream_t,
// execution.hpp:19-27), and link libhipcstone.so.
//
// Instantiations the reference has and this file leaves out, because hipcstone computes in one real type: the mixed
// (double coordinates, float h) computeBoundingBoxGpu of source_center_gpu.cu:100 and the (double, float, double)
// computeLeafSourceCenterGpu of :169.
#include <cstdint>
#include <span>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "cstone/focus/source_center_gpu.h"
#include "cstone/halos/gather_halos_gpu.h"
#include "cstone/primitives/primitives_gpu.h"
#include "cstone/sfc/sfc_gpu.h"
#include "cstone/traversal/collisions_gpu.h"
#include "cstone/tree/csarray_gpu.h"
#include "cstone/tree/octree_gpu.h"

#include "cstone_hipcstone.hpp"

namespace cstone
{
namespace
{
using hipcstone_bridge::check;
using hipcstone_bridge::stream;
using hipcstone_bridge::toC;

// one growable device workspace per host thread (the reference allocates its temporaries inside each call)
struct Workspace
{
    void* p     = nullptr;
    size_t size = 0;
    void* get(size_t bytes)
    {
        if (bytes > size)
        {
            if (p) check(hipcstone_device_free(p), "workspace");
            p = nullptr;
            check(hipcstone_device_alloc(&p, bytes), "workspace");
            size = bytes;
        }
        return p;
    }
    ~Workspace()
    {
        if (p) hipcstone_device_free(p);
    }
};
thread_local Workspace ws;

// whether the last computeNodeOpsGpu decision on this host thread was all ones (rebalanceTreeGpu reports the
// opposite as "changed")
template<class K>
int& lastConverged()
{
    thread_local int v = 1;
    return v;
}

template<class K>
constexpr bool is64 = sizeof(K) == 8;
template<class T>
constexpr bool isF32 = std::is_same_v<T, float>;

} // namespace

// ---- sfc/sfc_gpu.h:24-26 (sfc_gpu.cu:46-62): KeyType is MortonKey<K> or HilbertKey<K>; removeKey is honoured as in
// the reference kernel (sfc_gpu.cu:29)
template<class KeyType, class T>
void computeSfcKeys(execution::Gpu exec, const T* x, const T* y, const T* z, KeyType* keys, size_t numKeys,
                    const Box<T>& box)
{
    using K                = typename KeyType::ValueType;
    constexpr bool hilbert = std::is_same_v<KeyType, HilbertKey<K>>;
    auto* k                = reinterpret_cast<K*>(keys);
    auto cb                = toC(box);
    auto s                 = stream(exec);
    hipcstone_status st;
    if constexpr (isF32<T> && is64<K>)
        st = hilbert ? hipcstone_compute_sfc_keys_hilbert_u64_f32(s, x, y, z, k, numKeys, &cb, 1)
                     : hipcstone_compute_sfc_keys_morton_u64_f32(s, x, y, z, k, numKeys, &cb, 1);
    else if constexpr (isF32<T>)
        st = hilbert ? hipcstone_compute_sfc_keys_hilbert_u32_f32(s, x, y, z, k, numKeys, &cb, 1)
                     : hipcstone_compute_sfc_keys_morton_u32_f32(s, x, y, z, k, numKeys, &cb, 1);
    else if constexpr (is64<K>)
        st = hilbert ? hipcstone_compute_sfc_keys_hilbert_u64_f64(s, x, y, z, k, numKeys, &cb, 1)
                     : hipcstone_compute_sfc_keys_morton_u64_f64(s, x, y, z, k, numKeys, &cb, 1);
    else
        st = hilbert ? hipcstone_compute_sfc_keys_hilbert_u32_f64(s, x, y, z, k, numKeys, &cb, 1)
                     : hipcstone_compute_sfc_keys_morton_u32_f64(s, x, y, z, k, numKeys, &cb, 1);
    check(st, "computeSfcKeys");
}

#define HC_KEYS(KT, T)                                                                                                 \
    template void computeSfcKeys(execution::Gpu, const T*, const T*, const T*, KT*, size_t, const Box<T>&);
HC_KEYS(MortonKey<unsigned>, float)
HC_KEYS(MortonKey<unsigned>, double)
HC_KEYS(MortonKey<uint64_t>, float)
HC_KEYS(MortonKey<uint64_t>, double)
HC_KEYS(HilbertKey<unsigned>, float)
HC_KEYS(HilbertKey<unsigned>, double)
HC_KEYS(HilbertKey<uint64_t>, float)
HC_KEYS(HilbertKey<uint64_t>, double)
#undef HC_KEYS

// ---- primitives/primitives_gpu.h:119-130 (primitives_gpu.cu:310-367): stable sort by key with u32 values
template<class KeyType, class ValueType>
uint64_t sortByKeyTempStorage(uint64_t numElements)
{
    static_assert(sizeof(ValueType) == 4, "hipcstone sorts 32-bit values");
    if constexpr (is64<KeyType>) return hipcstone_sort_by_key_workspace_size_u64(numElements);
    else return hipcstone_sort_by_key_workspace_size_u32(numElements);
}

template<class KeyType, class ValueType>
void sortByKey(execution::Gpu exec, KeyType* first, KeyType* last, ValueType* values, KeyType* keyBuf,
               ValueType* valueBuf, void* tmp, uint64_t tmpBytes)
{
    static_assert(sizeof(ValueType) == 4, "hipcstone sorts 32-bit values");
    size_t n = size_t(last - first);
    auto v   = reinterpret_cast<uint32_t*>(values);
    auto vb  = reinterpret_cast<uint32_t*>(valueBuf);
    if constexpr (is64<KeyType>)
        check(hipcstone_sort_by_key_u64(stream(exec), reinterpret_cast<uint64_t*>(first), v, n,
                                        reinterpret_cast<uint64_t*>(keyBuf), vb, tmp, tmpBytes, 0),
              "sortByKey");
    else
        check(hipcstone_sort_by_key_u32(stream(exec), reinterpret_cast<uint32_t*>(first), v, n,
                                        reinterpret_cast<uint32_t*>(keyBuf), vb, tmp, tmpBytes, 0),
              "sortByKey");
}

#define HC_SORT(K, V)                                                                                                  \
    template uint64_t sortByKeyTempStorage<K, V>(uint64_t);                                                            \
    template void sortByKey(execution::Gpu, K*, K*, V*, K*, V*, void*, uint64_t);
HC_SORT(unsigned, unsigned)
HC_SORT(unsigned, int)
HC_SORT(uint64_t, unsigned)
HC_SORT(uint64_t, int)
#undef HC_SORT

// ---- tree/csarray_gpu.h:41-75 (csarray_gpu.cu:124-127, 207-210, 233-236)
template<class KeyType>
void computeNodeCountsGpu(execution::Gpu exec, const KeyType* tree, unsigned* counts, TreeNodeIndex numNodes,
                          std::span<const KeyType> keys, unsigned maxCount, bool useCountsAsGuess)
{
    if constexpr (is64<KeyType>)
        check(hipcstone_compute_node_counts_u64(stream(exec), tree, counts, numNodes, keys.data(), keys.size(),
                                                maxCount, useCountsAsGuess),
              "computeNodeCountsGpu");
    else
        check(hipcstone_compute_node_counts_u32(stream(exec), tree, counts, numNodes, keys.data(), keys.size(),
                                                maxCount, useCountsAsGuess),
              "computeNodeCountsGpu");
}

//! returns the new node count on the host, as the reference does (one synchronization)
template<class KeyType>
TreeNodeIndex computeNodeOpsGpu(execution::Gpu exec, const KeyType* tree, TreeNodeIndex numNodes,
                                const unsigned* counts, unsigned bucketSize, TreeNodeIndex* nodeOps)
{
    int32_t newNumNodes = 0;
    int converged       = 0;
    void* w             = ws.get(hipcstone_node_ops_workspace_size(numNodes));
    if constexpr (is64<KeyType>)
        check(hipcstone_compute_node_ops_u64(stream(exec), tree, numNodes, counts, bucketSize, nodeOps, w, ws.size,
                                             &newNumNodes, &converged),
              "computeNodeOpsGpu");
    else
        check(hipcstone_compute_node_ops_u32(stream(exec), tree, numNodes, counts, bucketSize, nodeOps, w, ws.size,
                                             &newNumNodes, &converged),
              "computeNodeOpsGpu");
    lastConverged<KeyType>() = converged;
    return newNumNodes;
}

//! returns whether the tree changed (the reference's change counter, csarray_gpu.cu:212-231): it did exactly when
//! the decision that produced nodeOps was not all ones
template<class KeyType>
bool rebalanceTreeGpu(execution::Gpu exec, const KeyType* tree, TreeNodeIndex numNodes, TreeNodeIndex newNumNodes,
                      const TreeNodeIndex* nodeOps, KeyType* newTree)
{
    if constexpr (is64<KeyType>)
        check(hipcstone_rebalance_tree_u64(stream(exec), tree, numNodes, newNumNodes, nodeOps, newTree),
              "rebalanceTreeGpu");
    else
        check(hipcstone_rebalance_tree_u32(stream(exec), tree, numNodes, newNumNodes, nodeOps, newTree),
              "rebalanceTreeGpu");
    return !lastConverged<KeyType>();
}

template void computeNodeCountsGpu(execution::Gpu, const unsigned*, unsigned*, TreeNodeIndex, std::span<const unsigned>,
                                   unsigned, bool);
template void computeNodeCountsGpu(execution::Gpu, const uint64_t*, unsigned*, TreeNodeIndex, std::span<const uint64_t>,
                                   unsigned, bool);
template TreeNodeIndex computeNodeOpsGpu(execution::Gpu, const unsigned*, TreeNodeIndex, const unsigned*, unsigned,
                                         TreeNodeIndex*);
template TreeNodeIndex computeNodeOpsGpu(execution::Gpu, const uint64_t*, TreeNodeIndex, const unsigned*, unsigned,
                                         TreeNodeIndex*);
template bool rebalanceTreeGpu(execution::Gpu, const unsigned*, TreeNodeIndex, TreeNodeIndex, const TreeNodeIndex*,
                               unsigned*);
template bool rebalanceTreeGpu(execution::Gpu, const uint64_t*, TreeNodeIndex, TreeNodeIndex, const TreeNodeIndex*,
                               uint64_t*);

// ---- tree/octree_gpu.h:35-45 (octree_gpu.cu:170-203): the OctreeView fields map one to one onto the C arguments;
// d_levelRange is the device level range, levelRange its host copy
template<class KeyType>
void buildOctreeGpu(execution::Gpu exec, const KeyType* cstoneTree, OctreeView<KeyType> d)
{
    size_t bytes;
    if constexpr (is64<KeyType>) bytes = hipcstone_build_octree_workspace_size_u64(d.numLeafNodes);
    else bytes = hipcstone_build_octree_workspace_size_u32(d.numLeafNodes);
    void* w = ws.get(bytes);
    if constexpr (is64<KeyType>)
        check(hipcstone_build_octree_u64(stream(exec), cstoneTree, d.numLeafNodes, d.prefixes, d.childOffsets,
                                         d.parents, d.d_levelRange, d.internalToLeaf, d.leafToInternal, w, ws.size,
                                         d.levelRange),
              "buildOctreeGpu");
    else
        check(hipcstone_build_octree_u32(stream(exec), cstoneTree, d.numLeafNodes, d.prefixes, d.childOffsets,
                                         d.parents, d.d_levelRange, d.internalToLeaf, d.leafToInternal, w, ws.size,
                                         d.levelRange),
              "buildOctreeGpu");
}

//! the overload with caller buffers: hipcstone's link needs one contiguous workspace, so the caller's CUB buffer is
//! used when it is large enough, else the thread's workspace
template<class KeyType>
void buildOctreeGpu(execution::Gpu exec, const KeyType* cstoneTree, OctreeView<KeyType> d,
                    std::span<KeyType> /*keyBuf*/, std::span<TreeNodeIndex> /*valueBuf*/, std::span<char> cubTmp)
{
    size_t bytes;
    if constexpr (is64<KeyType>) bytes = hipcstone_build_octree_workspace_size_u64(d.numLeafNodes);
    else bytes = hipcstone_build_octree_workspace_size_u32(d.numLeafNodes);
    void* w  = cubTmp.size() >= bytes ? static_cast<void*>(cubTmp.data()) : ws.get(bytes);
    size_t b = cubTmp.size() >= bytes ? cubTmp.size() : ws.size;
    if constexpr (is64<KeyType>)
        check(hipcstone_build_octree_u64(stream(exec), cstoneTree, d.numLeafNodes, d.prefixes, d.childOffsets,
                                         d.parents, d.d_levelRange, d.internalToLeaf, d.leafToInternal, w, b,
                                         d.levelRange),
              "buildOctreeGpu");
    else
        check(hipcstone_build_octree_u32(stream(exec), cstoneTree, d.numLeafNodes, d.prefixes, d.childOffsets,
                                         d.parents, d.d_levelRange, d.internalToLeaf, d.leafToInternal, w, b,
                                         d.levelRange),
              "buildOctreeGpu");
}

template void buildOctreeGpu(execution::Gpu, const uint32_t*, OctreeView<uint32_t>);
template void buildOctreeGpu(execution::Gpu, const uint64_t*, OctreeView<uint64_t>);
template void buildOctreeGpu(execution::Gpu, const uint32_t*, OctreeView<uint32_t>, std::span<uint32_t>,
                             std::span<TreeNodeIndex>, std::span<char>);
template void buildOctreeGpu(execution::Gpu, const uint64_t*, OctreeView<uint64_t>, std::span<uint64_t>,
                             std::span<TreeNodeIndex>, std::span<char>);

// ---- traversal/collisions_gpu.h:46-59 (collisions_gpu.cu:86-88). The reference signature carries no node count:
// INT32_MAX keeps the walk's structural guards only
template<class KeyType, class T>
void findHalosGpu(execution::Gpu exec, const KeyType* prefixes, const TreeNodeIndex* childOffsets,
                  const TreeNodeIndex* parents, const Vec3<T>* nodeCenters, const Vec3<T>* nodeSizes,
                  const KeyType* leaves, const Vec3<T>* searchCenters, const Vec3<T>* searchSizes, const Box<T>& box,
                  TreeNodeIndex firstNode, TreeNodeIndex lastNode, uint8_t* collisionFlags)
{
    auto c  = reinterpret_cast<const T*>(nodeCenters);
    auto sz = reinterpret_cast<const T*>(nodeSizes);
    auto sc = reinterpret_cast<const T*>(searchCenters);
    auto ss = reinterpret_cast<const T*>(searchSizes);
    auto cb = toC(box);
    auto s  = stream(exec);
    hipcstone_status st;
    if constexpr (is64<KeyType> && isF32<T>)
        st = hipcstone_find_halos_u64_f32(s, prefixes, childOffsets, parents, c, sz, leaves, sc, ss, &cb, firstNode,
                                          lastNode, INT32_MAX, collisionFlags, nullptr, 0);
    else if constexpr (is64<KeyType>)
        st = hipcstone_find_halos_u64_f64(s, prefixes, childOffsets, parents, c, sz, leaves, sc, ss, &cb, firstNode,
                                          lastNode, INT32_MAX, collisionFlags, nullptr, 0);
    else if constexpr (isF32<T>)
        st = hipcstone_find_halos_u32_f32(s, prefixes, childOffsets, parents, c, sz, leaves, sc, ss, &cb, firstNode,
                                          lastNode, INT32_MAX, collisionFlags, nullptr, 0);
    else
        st = hipcstone_find_halos_u32_f64(s, prefixes, childOffsets, parents, c, sz, leaves, sc, ss, &cb, firstNode,
                                          lastNode, INT32_MAX, collisionFlags, nullptr, 0);
    check(st, "findHalosGpu");
}

#define HC_HALOS(K, T)                                                                                                 \
    template void findHalosGpu(execution::Gpu, const K*, const TreeNodeIndex*, const TreeNodeIndex*, const Vec3<T>*,   \
                               const Vec3<T>*, const K*, const Vec3<T>*, const Vec3<T>*, const Box<T>&,               \
                               TreeNodeIndex, TreeNodeIndex, uint8_t*);
HC_HALOS(uint32_t, float)
HC_HALOS(uint64_t, float)
HC_HALOS(uint64_t, double)
#undef HC_HALOS

// ---- halos/gather_halos_gpu.h:24-31 (gather_halos_gpu.cu:60-74): elements are opaque records
template<class T, class IndexType>
void gatherRanges(execution::Gpu exec, const IndexType* rangeScan, const IndexType* rangeOffsets, int numRanges,
                  const T* src, T* buffer, size_t bufferSize)
{
    static_assert(sizeof(IndexType) == 4, "range offsets are 32-bit");
    check(hipcstone_gather_ranges(stream(exec), reinterpret_cast<const uint32_t*>(rangeScan),
                                  reinterpret_cast<const uint32_t*>(rangeOffsets), numRanges, src, buffer, bufferSize,
                                  sizeof(T)),
          "gatherRanges");
}

template void gatherRanges(execution::Gpu, const unsigned*, const unsigned*, int, const int*, int*, size_t);
template void gatherRanges(execution::Gpu, const unsigned*, const unsigned*, int, const util::array<float, 1>*,
                           util::array<float, 1>*, size_t);
template void gatherRanges(execution::Gpu, const unsigned*, const unsigned*, int, const util::array<float, 2>*,
                           util::array<float, 2>*, size_t);
template void gatherRanges(execution::Gpu, const unsigned*, const unsigned*, int, const util::array<float, 3>*,
                           util::array<float, 3>*, size_t);
template void gatherRanges(execution::Gpu, const unsigned*, const unsigned*, int, const util::array<float, 4>*,
                           util::array<float, 4>*, size_t);

// ---- focus/source_center_gpu.h:40-116 (source_center_gpu.cu:99-101, 168-170, 207-210, 242-245, 273-276, 307-310)
template<class Tc, class Th>
void computeBoundingBoxGpu(execution::Gpu exec, const Tc* x, const Tc* y, const Tc* z, const Th* h,
                           const LocalIndex* layout, TreeNodeIndex first, TreeNodeIndex last, Th scale,
                           Vec3<Tc>* centers, Vec3<Tc>* sizes)
{
    static_assert(std::is_same_v<Tc, Th>, "hipcstone computes search boxes in one real type");
    auto c = reinterpret_cast<Tc*>(centers);
    auto s = reinterpret_cast<Tc*>(sizes);
    if constexpr (isF32<Tc>)
        check(hipcstone_compute_bounding_box_f32(stream(exec), x, y, z, h, layout, first, last, scale, c, s),
              "computeBoundingBoxGpu");
    else
        check(hipcstone_compute_bounding_box_f64(stream(exec), x, y, z, h, layout, first, last, scale, c, s),
              "computeBoundingBoxGpu");
}

template void computeBoundingBoxGpu(execution::Gpu, const double*, const double*, const double*, const double*,
                                    const LocalIndex*, TreeNodeIndex, TreeNodeIndex, double, Vec3<double>*,
                                    Vec3<double>*);
template void computeBoundingBoxGpu(execution::Gpu, const float*, const float*, const float*, const float*,
                                    const LocalIndex*, TreeNodeIndex, TreeNodeIndex, float, Vec3<float>*,
                                    Vec3<float>*);

template<class Tc, class Tm, class Tf>
void computeLeafSourceCenterGpu(execution::Gpu exec, const Tc* x, const Tc* y, const Tc* z, const Tm* m,
                                const TreeNodeIndex* leafToInternal, TreeNodeIndex numLeaves, const LocalIndex* layout,
                                Vec4<Tf>* centers)
{
    static_assert(std::is_same_v<Tc, Tm> && std::is_same_v<Tc, Tf>, "hipcstone sums in one real type");
    if constexpr (isF32<Tc>)
        check(hipcstone_leaf_mass_centers_f32(stream(exec), x, y, z, m, leafToInternal, numLeaves, layout,
                                              reinterpret_cast<float*>(centers)),
              "computeLeafSourceCenterGpu");
    else
        check(hipcstone_leaf_mass_centers_f64(stream(exec), x, y, z, m, leafToInternal, numLeaves, layout,
                                              reinterpret_cast<double*>(centers)),
              "computeLeafSourceCenterGpu");
}

template void computeLeafSourceCenterGpu(execution::Gpu, const double*, const double*, const double*, const double*,
                                         const TreeNodeIndex*, TreeNodeIndex, const LocalIndex*, Vec4<double>*);
template void computeLeafSourceCenterGpu(execution::Gpu, const float*, const float*, const float*, const float*,
                                         const TreeNodeIndex*, TreeNodeIndex, const LocalIndex*, Vec4<float>*);

//! levelRange is a host array here, as in the reference's caller (octree_focus_mpi.hpp:395-400)
template<class T>
void upsweepCentersGpu(execution::Gpu exec, int numLevels, const TreeNodeIndex* levelRange,
                       const TreeNodeIndex* childOffsets, SourceCenterType<T>* centers)
{
    if constexpr (isF32<T>)
        check(hipcstone_upsweep_centers_f32(stream(exec), numLevels, levelRange, childOffsets,
                                            reinterpret_cast<float*>(centers)),
              "upsweepCentersGpu");
    else
        check(hipcstone_upsweep_centers_f64(stream(exec), numLevels, levelRange, childOffsets,
                                            reinterpret_cast<double*>(centers)),
              "upsweepCentersGpu");
}

template void upsweepCentersGpu(execution::Gpu, int, const TreeNodeIndex*, const TreeNodeIndex*,
                                SourceCenterType<float>*);
template void upsweepCentersGpu(execution::Gpu, int, const TreeNodeIndex*, const TreeNodeIndex*,
                                SourceCenterType<double>*);

template<class KeyType, class T>
void computeGeoCentersGpu(execution::Gpu exec, const KeyType* prefixes, TreeNodeIndex numNodes, Vec3<T>* centers,
                          Vec3<T>* sizes, const Box<T>& box)
{
    auto cb = toC(box);
    auto c  = reinterpret_cast<T*>(centers);
    auto s  = reinterpret_cast<T*>(sizes);
    hipcstone_status st;
    if constexpr (is64<KeyType> && isF32<T>)
        st = hipcstone_compute_geo_centers_u64_f32(stream(exec), prefixes, numNodes, c, s, &cb);
    else if constexpr (is64<KeyType>)
        st = hipcstone_compute_geo_centers_u64_f64(stream(exec), prefixes, numNodes, c, s, &cb);
    else if constexpr (isF32<T>)
        st = hipcstone_compute_geo_centers_u32_f32(stream(exec), prefixes, numNodes, c, s, &cb);
    else
        st = hipcstone_compute_geo_centers_u32_f64(stream(exec), prefixes, numNodes, c, s, &cb);
    check(st, "computeGeoCentersGpu");
}

template<class KeyType, class T>
void geoMacSpheresGpu(execution::Gpu exec, const KeyType* prefixes, TreeNodeIndex numNodes,
                      SourceCenterType<T>* centers, float invTheta, const Box<T>& box)
{
    auto cb = toC(box);
    auto c  = reinterpret_cast<T*>(centers);
    hipcstone_status st;
    if constexpr (is64<KeyType> && isF32<T>)
        st = hipcstone_geo_mac_spheres_u64_f32(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    else if constexpr (is64<KeyType>)
        st = hipcstone_geo_mac_spheres_u64_f64(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    else if constexpr (isF32<T>)
        st = hipcstone_geo_mac_spheres_u32_f32(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    else
        st = hipcstone_geo_mac_spheres_u32_f64(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    check(st, "geoMacSpheresGpu");
}

template<class KeyType, class T>
void setMacGpu(execution::Gpu exec, const KeyType* prefixes, TreeNodeIndex numNodes, Vec4<T>* macSpheres,
               float invTheta, const Box<T>& box)
{
    auto cb = toC(box);
    auto c  = reinterpret_cast<T*>(macSpheres);
    hipcstone_status st;
    if constexpr (is64<KeyType> && isF32<T>)
        st = hipcstone_set_mac_u64_f32(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    else if constexpr (is64<KeyType>)
        st = hipcstone_set_mac_u64_f64(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    else if constexpr (isF32<T>)
        st = hipcstone_set_mac_u32_f32(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    else
        st = hipcstone_set_mac_u32_f64(stream(exec), prefixes, numNodes, c, invTheta, &cb);
    check(st, "setMacGpu");
}

#define HC_GEO(K, T)                                                                                                   \
    template void computeGeoCentersGpu(execution::Gpu, const K*, TreeNodeIndex, Vec3<T>*, Vec3<T>*, const Box<T>&);   \
    template void geoMacSpheresGpu(execution::Gpu, const K*, TreeNodeIndex, SourceCenterType<T>*, float,             \
                                   const Box<T>&);                                                                     \
    template void setMacGpu(execution::Gpu, const K*, TreeNodeIndex, Vec4<T>*, float, const Box<T>&);
HC_GEO(uint32_t, float)
HC_GEO(uint32_t, double)
HC_GEO(uint64_t, float)
HC_GEO(uint64_t, double)
#undef HC_GEO

// findNeighbors(execution::Gpu, ...) of cstone_hipcstone.hpp, instantiated for the reference's CPU sets
template void findNeighbors(execution::Gpu, const float*, const float*, const float*, const float*, LocalIndex,
                            LocalIndex, const Box<float>&, const OctreeNsView<float, uint32_t>&, unsigned, LocalIndex*,
                            unsigned*);
template void findNeighbors(execution::Gpu, const float*, const float*, const float*, const float*, LocalIndex,
                            LocalIndex, const Box<float>&, const OctreeNsView<float, uint64_t>&, unsigned, LocalIndex*,
                            unsigned*);
template void findNeighbors(execution::Gpu, const double*, const double*, const double*, const double*, LocalIndex,
                            LocalIndex, const Box<double>&, const OctreeNsView<double, uint32_t>&, unsigned,
                            LocalIndex*, unsigned*);
template void findNeighbors(execution::Gpu, const double*, const double*, const double*, const double*, LocalIndex,
                            LocalIndex, const Box<double>&, const OctreeNsView<double, uint64_t>&, unsigned,
                            LocalIndex*, unsigned*);

} // namespace cstone
