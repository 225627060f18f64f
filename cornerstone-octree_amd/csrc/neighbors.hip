// a16: radius neighbor search with the per-particle semantics of the reference CPU path (findneighbors.hpp:77-177,
// traversal.hpp:26-69): the library's launches of the wave walk in hipcstone_nbwalk.hpp (which documents the
// semantics and the design), row-major or slot-major, plain or segmented (the batched clouds).
#include <limits>
#include <type_traits>

#include "common.hpp"
#include "hipcstone_nbwalk.hpp"
#include "runtime.hpp"

namespace hipcstone
{

// SlotMajor: the wave's 64 rows are stored slot-major (neighbor k of the wave's lane l at k * 64 + l) instead of
// row-major (l * ngmax + k), so that a consumer with one thread per target reads each slot with one coalesced load.
// Segmented: many independent trees in one set of arrays (the batched clouds). Wave `group` then takes its targets
// [first, last) (at most 64, inside one cloud) and the index of its cloud's root node from the wave table
// (waves[3 g], waves[3 g + 1], cloud = waves[3 g + 2], root = roots[cloud]); rows and counts are indexed by the
// particle index itself (firstId = 0), and a child block may lie before its parent (the batched link places children
// by internal-node slot, not level by level), so only the bounds and depth guards and the step budget apply.
// Waves per workgroup: 4 by default; 2 and 8 are the other shapes the autotuner measures
// (hipcstone_set_neighbor_waves). Every wave owns its own targets, stack and leaf list, so the shape changes only
// how waves are grouped onto CUs (LDS per workgroup, scheduling granularity), never a result.
// The fp32 open-box walk fits 8 waves per SIMD when its scalar registers are capped (the compiler then keeps some
// wave-uniform values of the walk in vector-register lanes): more waves to cover the latency of its LDS, scalar-load
// and compare-to-ballot dependencies, measured 6 % faster than the 7 waves its uncapped allocation allows. The other
// variants keep the compiler's own allocation.
template<class T, bool AnyPbc>
constexpr int kNbMinWavesPerEu = std::is_same_v<T, float> && !AnyPbc ? 8 : 1;

template<class T, bool AnyPbc, bool SlotMajor, bool Segmented, int kNbWaves>
__global__ __launch_bounds__(kNbWaves * kWave) __attribute__((amdgpu_waves_per_eu(kNbMinWavesPerEu<T, AnyPbc>, 8)))
void findNeighborsKernel(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, const T* __restrict__ h,
    uint32_t firstId, uint32_t lastId, typename BoxOf<T>::type box, const int32_t* __restrict__ childOffsets,
    const int32_t* __restrict__ internalToLeaf, const uint32_t* __restrict__ layout, const T* __restrict__ centers,
    const T* __restrict__ sizes, int32_t numNodes, float ext, uint32_t ngmax, uint32_t* __restrict__ neighbors,
    uint32_t* __restrict__ neighborsCount, uint32_t numBlocksPadded, uint32_t* __restrict__ faults,
    const uint32_t* __restrict__ waves = nullptr, uint32_t numWaves = 0, const int32_t* __restrict__ roots = nullptr,
    const nbwalk::ChildGroup* __restrict__ groups = nullptr)
{
    __shared__ nbwalk::Shared<T, kNbWaves> sh;
    const int lane       = laneId();
    const int w          = threadIdx.x / kWave;
    const uint32_t block = xcdContiguousBlock(blockIdx.x, numBlocksPadded);
    const uint32_t group = uniform(block * uint32_t(kNbWaves) + uint32_t(w));
    uint64_t first;
    int32_t root = 0;
    if constexpr (Segmented)
    {
        if (group >= numWaves) return;
        first  = uniform(waves[3 * group]);
        lastId = uniform(waves[3 * group + 1]);
        root   = uniform(roots[uniform(waves[3 * group + 2])]);
        if (root < 0 || root >= numNodes)
        {
            if (lane == 0) atomicOr(faults, HIPCSTONE_FAULT_MALFORMED_TREE);
            return;
        }
    }
    else
    {
        first = uint64_t(firstId) + uint64_t(group) * kWave;
        if (first >= lastId) return;
    }
    const uint32_t i     = uint32_t(first) + uint32_t(lane);
    const bool valid     = i < lastId;
    const size_t rowBase = Segmented ? size_t(first) : size_t(group) * kWave; // row of lane 0 in the output

    // row-major: the rows of the wave's valid targets; slot-major: all 64 rows of the wave
    const uint32_t validRows = uint64_t(lastId) - first < uint64_t(kWave) ? uint32_t(uint64_t(lastId) - first) : kWave;
    const uint32_t rowsBytes = ngmax ? (SlotMajor ? uint32_t(kWave) : validRows) * ngmax * uint32_t(sizeof(uint32_t))
                                     : 0u;
    const nbwalk::Tree<T> tree{childOffsets, internalToLeaf, layout, centers, sizes, numNodes, ext, groups};
    const uint32_t count = nbwalk::walk<T, AnyPbc, SlotMajor, Segmented, kNbWaves>(
        x, y, z, h, box, tree, root, first, valid, ngmax, ngmax ? neighbors + rowBase * ngmax : nullptr, rowsBytes,
        sh, w, faults);
    if (valid) neighborsCount[i - firstId] = count;
}

// the segmented search of the batched clouds (row-major rows, open boxes): see findNeighborsKernel
template<class T>
hipcstone_status findNeighborsSegmented(hipStream_t s, const T* x, const T* y, const T* z, const T* h,
                                        const typename BoxOf<T>::type& box, const int32_t* childOffsets,
                                        const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                        const T* sizes, int32_t numNodesCap, uint32_t ngmax, uint32_t* neighbors,
                                        uint32_t* neighborsCount, const uint32_t* waves, uint32_t numWaves,
                                        const int32_t* roots, const nbwalk::ChildGroup* groups)
{
    if (numWaves == 0) return HIPCSTONE_OK;
    if (ngmax > (1u << 20)) return invalid("findNeighbors: ngmax above 2^20");
    uint32_t* faults = faultWord();
    if (!faults) return HIPCSTONE_HIP_ERROR;
    auto launch = [&](auto wv) {
        constexpr int W = decltype(wv)::value;
        unsigned blocks = ceilDiv(size_t(numWaves), W);
        blocks          = (blocks + kXcds - 1) / kXcds * kXcds;
        findNeighborsKernel<T, false, false, true, W><<<blocks, W * kWave, 0, s>>>(
            x, y, z, h, 0u, 0u, box, childOffsets, internalToLeaf, layout, centers, sizes, numNodesCap, 1.0f, ngmax,
            neighbors, neighborsCount, blocks, faults, waves, numWaves, roots, groups);
    };
    switch (neighborWaves())
    {
    case 2: launch(std::integral_constant<int, 2>{}); break;
    case 8: launch(std::integral_constant<int, 8>{}); break;
    default: launch(std::integral_constant<int, 4>{}); break;
    }
    HC_CHECK_LAUNCH("findNeighborsKernel (segmented)");
    return HIPCSTONE_OK;
}

template hipcstone_status findNeighborsSegmented<float>(hipStream_t, const float*, const float*, const float*,
                                                        const float*, const BoxOf<float>::type&, const int32_t*,
                                                        const int32_t*, const uint32_t*, const float*, const float*,
                                                        int32_t, uint32_t, uint32_t*, uint32_t*, const uint32_t*,
                                                        uint32_t, const int32_t*, const nbwalk::ChildGroup*);

size_t neighborGroupsBytes(int32_t numNodes) { return nbwalk::childGroupCount(numNodes) * sizeof(nbwalk::ChildGroup); }

hipcstone_status buildChildGroups(hipStream_t s, const float* centers, const float* sizes,
                                  const int32_t* childOffsets, const int32_t* internalToLeaf, int32_t numNodes,
                                  nbwalk::ChildGroup* groups)
{
    const uint32_t numGroups = uint32_t(nbwalk::childGroupCount(numNodes));
    if (numGroups == 0) return HIPCSTONE_OK;
    nbwalk::childGroupsKernel<<<(numGroups + 255) / 256, 256, 0, s>>>(centers, sizes, childOffsets, internalToLeaf,
                                                                        numGroups, groups);
    HC_CHECK_LAUNCH("childGroupsKernel");
    return HIPCSTONE_OK;
}

template<class T>
static hipcstone_status findNeighbors(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h,
                                      uint32_t firstId, uint32_t lastId, const typename BoxOf<T>::type* box,
                                      const int32_t* childOffsets, const int32_t* parents,
                                      const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                      const T* sizes, int32_t numNodes, float ext, uint32_t ngmax,
                                      uint32_t* neighbors, uint32_t* neighborsCount, bool slotMajor,
                                      void* groupsWs = nullptr, size_t groupsWsBytes = 0)
{
    if (lastId <= firstId) return HIPCSTONE_OK;
    if (!x || !y || !z || !h || !box || !childOffsets || !parents || !internalToLeaf || !layout || !centers ||
        !sizes || !neighborsCount)
        return invalid("findNeighbors: null pointer");
    if (ngmax > 0 && !neighbors) return invalid("findNeighbors: neighbors is null with ngmax > 0");
    if (ngmax > (1u << 20)) return invalid("findNeighbors: ngmax above 2^20");
    if (numNodes <= 0) return invalid("findNeighbors: numNodes must be positive");
    size_t waveGroups = (size_t(lastId - firstId) + kWave - 1) / kWave;
    uint32_t* faults = faultWord();
    if (!faults) return HIPCSTONE_HIP_ERROR;
    bool anyPbc     = box->bnd[0] == HIPCSTONE_PERIODIC || box->bnd[1] == HIPCSTONE_PERIODIC ||
                  box->bnd[2] == HIPCSTONE_PERIODIC;
    // fp32 open boxes walk through the compact child blocks, built here into the caller's workspace (or into memory
    // allocated on the stream for this call)
    void* owned                       = nullptr;
    const nbwalk::ChildGroup* groups = nullptr;
    if constexpr (std::is_same_v<T, float>)
    {
        const size_t need = neighborGroupsBytes(numNodes);
        if (!anyPbc && need > 0)
        {
            void* dst = groupsWs;
            if (!dst || groupsWsBytes < need)
            {
                HC_CHECK(hipMallocAsync(&owned, need, stream(s)));
                dst = owned;
            }
            if (auto e = buildChildGroups(stream(s), centers, sizes, childOffsets, internalToLeaf, numNodes,
                                          static_cast<nbwalk::ChildGroup*>(dst)))
            {
                if (owned) (void)hipFreeAsync(owned, stream(s));
                return e;
            }
            groups = static_cast<const nbwalk::ChildGroup*>(dst);
        }
    }
    auto shaped = [&](auto pbc, auto slots, auto wv) {
        constexpr int W = decltype(wv)::value;
        unsigned blocks = ceilDiv(waveGroups, W);
        blocks          = (blocks + kXcds - 1) / kXcds * kXcds; // whole rounds over the XCDs; extra groups exit
        findNeighborsKernel<T, decltype(pbc)::value, decltype(slots)::value, false, W>
            <<<blocks, W * kWave, 0, stream(s)>>>(x, y, z, h, firstId, lastId, *box, childOffsets, internalToLeaf,
                                                  layout, centers, sizes, numNodes, ext, ngmax, neighbors,
                                                  neighborsCount, blocks, faults, nullptr, 0u, nullptr,
                                                  decltype(pbc)::value ? nullptr : groups);
    };
    auto launch = [&](auto pbc, auto slots) {
        switch (neighborWaves())
        {
        case 2: shaped(pbc, slots, std::integral_constant<int, 2>{}); break;
        case 8: shaped(pbc, slots, std::integral_constant<int, 8>{}); break;
        default: shaped(pbc, slots, std::integral_constant<int, 4>{}); break;
        }
    };
    if (anyPbc && slotMajor) launch(std::true_type{}, std::true_type{});
    else if (anyPbc) launch(std::true_type{}, std::false_type{});
    else if (slotMajor) launch(std::false_type{}, std::true_type{});
    else launch(std::false_type{}, std::false_type{});
    hipError_t launched = hipGetLastError();
    if (owned) HC_CHECK(hipFreeAsync(owned, stream(s)));
    if (launched != hipSuccess) return hipFail(launched, "findNeighborsKernel");
    return HIPCSTONE_OK;
}

template<class T>
hipcstone_status findNeighborsWs(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t firstId,
                                 uint32_t lastId, const typename BoxOf<T>::type* box, const int32_t* childOffsets,
                                 const int32_t* parents, const int32_t* internalToLeaf, const uint32_t* layout,
                                 const T* centers, const T* sizes, int32_t numNodes, float ext, uint32_t ngmax,
                                 uint32_t* neighbors, uint32_t* neighborsCount, void* groupsWs, size_t groupsWsBytes)
{
    return findNeighbors<T>(s, x, y, z, h, firstId, lastId, box, childOffsets, parents, internalToLeaf, layout,
                            centers, sizes, numNodes, ext, ngmax, neighbors, neighborsCount, false, groupsWs,
                            groupsWsBytes);
}

#define HIPCSTONE_INST_NEIGHBORS_WS(T)                                                                                 \
    template hipcstone_status findNeighborsWs<T>(                                                                      \
        hipcstone_stream, const T*, const T*, const T*, const T*, uint32_t, uint32_t, const BoxOf<T>::type*,           \
        const int32_t*, const int32_t*, const int32_t*, const uint32_t*, const T*, const T*, int32_t, float, uint32_t, \
        uint32_t*, uint32_t*, void*, size_t);
HIPCSTONE_INST_NEIGHBORS_WS(float)
HIPCSTONE_INST_NEIGHBORS_WS(double)

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

#define HIPCSTONE_DEF_NEIGHBORS(RS, T)                                                                                 \
    hipcstone_status hipcstone_find_neighbors_##RS(                                                                    \
        hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t first_id, uint32_t last_id,      \
        const hipcstone_box_##RS* box, const int32_t* child_offsets, const int32_t* parents,                          \
        const int32_t* internal_to_leaf, const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes,  \
        float ext, uint32_t ngmax, uint32_t* neighbors, uint32_t* neighbors_count)                                     \
    {                                                                                                                  \
        return findNeighbors<T>(s, x, y, z, h, first_id, last_id, box, child_offsets, parents, internal_to_leaf,      \
                                layout, centers, sizes, num_nodes, ext, ngmax, neighbors, neighbors_count, false);     \
    }                                                                                                                  \
    hipcstone_status hipcstone_find_neighbors_slot_major_##RS(                                                         \
        hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t first_id, uint32_t last_id,      \
        const hipcstone_box_##RS* box, const int32_t* child_offsets, const int32_t* parents,                          \
        const int32_t* internal_to_leaf, const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes,  \
        float ext, uint32_t ngmax, uint32_t* neighbors, uint32_t* neighbors_count)                                     \
    {                                                                                                                  \
        return findNeighbors<T>(s, x, y, z, h, first_id, last_id, box, child_offsets, parents, internal_to_leaf,      \
                                layout, centers, sizes, num_nodes, ext, ngmax, neighbors, neighbors_count, true);      \
    }

HIPCSTONE_DEF_NEIGHBORS(f32, float)
HIPCSTONE_DEF_NEIGHBORS(f64, double)

} // extern "C"
