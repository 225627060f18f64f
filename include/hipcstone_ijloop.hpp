/*
 * hipcstone ijloop: per-particle interaction loops over radius neighbors, for HIP C++ callers.
 *
 * Replaces cstone::ijloop with the always-traverse neighborhood (traversal/ijloop/ijloop.hpp:29-146,
 * common.hpp:29-258, gpu_alwaystraverse.cuh:31-208): a builder { ngmax }.build(stream, tree, box, totalBodies,
 * groups, x, y, z, h) returns a neighborhood whose ijLoop(input, output, interaction, postamble) computes, for every
 * body i of the groups,
 *
 *     result = interaction(iData, iData, {0,0,0}, 0)
 *     for each of the first ngmax neighbors j of i, ascending j:   result ⊕= interaction(iData, jData, r_ij, |r_ij|²)
 *     output[i] = postamble(iData, unwrapped result)
 *
 * with iData = (i, (x,y,z)_i, h_i, input_i...), r_ij = p_i − p_j (minimum image for bodies whose 2h box leaves a
 * periodic domain), |r|² = r0² + (r1² + r2²), ⊕ = + by default, min / max for reduction::min / max, and the
 * symmetric::even / odd markers unwrapped (they only matter to symmetric neighborhoods). The neighbor lists are the
 * ones hipcstone_find_neighbors_* produces (the reference CPU findNeighbors lists), each body's contributions are
 * folded sequentially in list order, so results equal the reference CPU ijloop's (cpu_alwaystraverse.hpp:90-113)
 * bit for bit when the interaction is evaluated with the same operations.
 *
 * The neighbor search and the fold run in one kernel, instantiated here, in the caller's translation unit, for the
 * caller's interaction and postamble (device-callable functors or lambdas), as the reference's GPU loop folds each
 * warp's lists right after that warp's traversal (gpu_alwaystraverse.cuh:32-91): a persistent grid of waves takes
 * groups of 64 consecutive bodies; each wave walks the tree for its group with the library's own walk
 * (hipcstone_nbwalk.hpp), writing the 64 lists into its own slice of a small scratch (64 * ngmax slots per resident
 * wave, reused group after group: the lists are never held in memory whole), then every lane folds its own body's
 * list from that slice, in list order. x, y, z, h of all totalBodies bodies are also staged as one
 * record per body at each ijLoop, so a neighbor's position and h come in one wide load (the values are copies, the
 * results do not change).
 * h must be an array (the reference also accepts a constant). Errors are thrown as std::runtime_error carrying
 * hipcstone_last_error().
 */
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>

#include "hipcstone.h"
#include "hipcstone_nbwalk.hpp"

namespace hipcstone::ijloop
{

using LocalIndex = uint32_t;

template<class T>
struct Vec3
{
    T v[3];
    __host__ __device__ constexpr T& operator[](int i) { return v[i]; }
    __host__ __device__ constexpr const T& operator[](int i) const { return v[i]; }
};

template<class T>
using Box = std::conditional_t<std::is_same_v<T, float>, hipcstone_box_f32, hipcstone_box_f64>;

// ---- result modifiers (ijloop.hpp:29-98) ----
namespace symmetric
{
template<class T>
struct even
{
    T value = {};
};
template<class T>
struct odd
{
    T value = {};
};
} // namespace symmetric

namespace reduction
{
template<class T>
struct min
{
    T value = std::numeric_limits<T>::max();
};
template<class T>
struct max
{
    T value = std::numeric_limits<T>::lowest();
};
} // namespace reduction

namespace detail
{
struct EmptyPostamble
{
    template<class ParticleData, class Result>
    __host__ __device__ constexpr Result operator()(ParticleData const&, Result const& result) const
    {
        return result;
    }
};
} // namespace detail

inline constexpr detail::EmptyPostamble empty_postamble;

struct Statistics
{
    const std::size_t numBodies, numBytes;
};

// traversal/groups.hpp:28-34
struct GroupView
{
    LocalIndex firstBody, lastBody;
    LocalIndex numGroups;
    const LocalIndex* groupStart;
    const LocalIndex* groupEnd;
};

// tree/octree.hpp:259-283 (device pointers of a linked octree and its leaf layout)
template<class T, class KeyType>
struct OctreeNsView
{
    int32_t numLeafNodes;
    int32_t numNodes;
    const KeyType* prefixes;
    const int32_t* childOffsets;
    const int32_t* parents;
    const int32_t* internalToLeaf;
    const int32_t* leafToInternal;
    const int32_t* levelRange;
    const KeyType* leaves;
    const LocalIndex* layout;
    const Vec3<T>* centers;
    const Vec3<T>* sizes;
    float searchExtFactor{1.0f};
};

namespace detail
{

template<class F, class Tuple, std::size_t... I>
__host__ __device__ constexpr void forEachIndex(F&& f, Tuple&& t, std::index_sequence<I...>)
{
    (f(std::get<I>(t), std::integral_constant<std::size_t, I>{}), ...);
}

template<class T>
__host__ __device__ constexpr void updateOne(T& r, T const& v)
{
    r += v;
}
template<class T>
__host__ __device__ constexpr void updateOne(reduction::min<T>& r, reduction::min<T> const& v)
{
    r.value = (v.value < r.value) ? v.value : r.value; // std::min(r, v)
}
template<class T>
__host__ __device__ constexpr void updateOne(reduction::max<T>& r, reduction::max<T> const& v)
{
    r.value = (r.value < v.value) ? v.value : r.value; // std::max(r, v)
}
template<class T>
__host__ __device__ constexpr void updateOne(symmetric::even<T>& r, symmetric::even<T> const& v)
{
    updateOne(r.value, v.value);
}
template<class T>
__host__ __device__ constexpr void updateOne(symmetric::odd<T>& r, symmetric::odd<T> const& v)
{
    updateOne(r.value, v.value);
}

template<class T>
__host__ __device__ constexpr T unwrapOne(T const& v)
{
    return v;
}
template<class T>
__host__ __device__ constexpr auto unwrapOne(reduction::min<T> const& v)
{
    return unwrapOne(v.value);
}
template<class T>
__host__ __device__ constexpr auto unwrapOne(reduction::max<T> const& v)
{
    return unwrapOne(v.value);
}
template<class T>
__host__ __device__ constexpr auto unwrapOne(symmetric::even<T> const& v)
{
    return unwrapOne(v.value);
}
template<class T>
__host__ __device__ constexpr auto unwrapOne(symmetric::odd<T> const& v)
{
    return unwrapOne(v.value);
}

template<class... Ts, std::size_t... I>
__host__ __device__ constexpr void updateResult(std::tuple<Ts...>& r, std::tuple<Ts...> const& v,
                                                std::index_sequence<I...>)
{
    (updateOne(std::get<I>(r), std::get<I>(v)), ...);
}

template<class... Ts, std::size_t... I>
__host__ __device__ constexpr auto unwrapModifiers(std::tuple<Ts...> const& v, std::index_sequence<I...>)
{
    return std::make_tuple(unwrapOne(std::get<I>(v))...);
}

// x, y, z, h of one body, staged as one record so that a neighbor's position and h come in one wide load
template<class T>
struct alignas(4 * sizeof(T)) Record
{
    T x, y, z, h;
};

template<class T>
__global__ __launch_bounds__(256) void stageRecordsKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                          const T* __restrict__ z, const T* __restrict__ h,
                                                          LocalIndex n, Record<T>* __restrict__ rec)
{
    LocalIndex i = LocalIndex(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) rec[i] = Record<T>{x[i], y[i], z[i], h[i]};
}

// (index, position, h, inputs...) of body i (common.hpp:39-55), from the staged records
template<class T, class... Ts, std::size_t... I>
__device__ inline auto loadParticleData(const Record<T>* rec, std::tuple<const Ts*...> const& input, LocalIndex i,
                                        std::index_sequence<I...>)
{
    const Record<T> r = rec[i];
    return std::tuple<LocalIndex, Vec3<T>, T, Ts...>(i, Vec3<T>{{r.x, r.y, r.z}}, r.h, std::get<I>(input)[i]...);
}


template<class... Ts, class... Vs, std::size_t... I>
__device__ inline void storeParticleData(std::tuple<Ts*...> const& output, LocalIndex i,
                                         std::tuple<Vs...> const& value, std::index_sequence<I...>)
{
    ((std::get<I>(output)[i] = std::get<I>(value)), ...);
}

// requiresPbcHandling (common.hpp:80-89): some dimension periodic and the body's 2h box leaves the domain
// (insideBox, boxoverlap.hpp:217-227)
template<class T>
__device__ inline bool requiresPbc(const Box<T>& box, const Vec3<T>& p, T hi)
{
    T s = T(2) * hi;
    bool inside = true;
    for (int d = 0; d < 3; ++d)
        inside = inside && (p[d] - s) >= box.lim[2 * d] && (p[d] + s) <= box.lim[2 * d + 1];
    return !inside;
}

// posDiffAndDistSq (common.hpp:91-107): r = p_i - p_j, minimum image per periodic dimension, |r|^2 right fold
template<class T>
__device__ inline void posDiffAndDistSq(bool usePbc, const Box<T>& box, const Vec3<T>& pi, const Vec3<T>& pj,
                                        Vec3<T>& r, T& d2)
{
    r[0] = pi[0] - pj[0];
    r[1] = pi[1] - pj[1];
    r[2] = pi[2] - pj[2];
    if (usePbc)
    {
        for (int d = 0; d < 3; ++d)
            r[d] -= T(box.bnd[d] == HIPCSTONE_PERIODIC) * box.len[d] * rint(r[d] * box.ilen[d]);
    }
    d2 = r[0] * r[0] + (r[1] * r[1] + r[2] * r[2]);
}

// body i belongs to the subgroups: the ranges [groupStart[g], groupEnd[g]) are ascending and disjoint
__device__ inline bool inGroups(const GroupView& groups, LocalIndex i)
{
    if (groups.numGroups == 0) return true;
    LocalIndex lo = 0, hi = groups.numGroups; // the first range that starts after i
    while (lo < hi)
    {
        LocalIndex mid = lo + (hi - lo) / 2;
        if (groups.groupStart[mid] <= i) lo = mid + 1;
        else hi = mid;
    }
    return lo > 0 && i < groups.groupEnd[lo - 1];
}

constexpr int kFusedWaves     = 4; // waves per workgroup of the fused kernel
constexpr unsigned kFoldBatch = 8; // neighbors whose data a fold step gathers at once

__global__ void zeroCountersKernel(uint32_t* counters)
{
    if (threadIdx.x < hipcstone::kXcds) counters[threadIdx.x] = 0;
}

/*
 * The fused always-traverse loop. Bodies [first, last) form groups of 64; the groups are split into 8 contiguous
 * ranges, one per XCD (workgroups are dealt to the XCDs round-robin, so workgroup b runs on XCD b % 8 and neighbouring
 * groups share tree nodes and particles in that XCD's L2). Each wave takes the next group of its XCD's range from
 * counters[xcd] and, when that range is exhausted, helps with the next XCD's. Per group: the walk stores the 64 lists
 * (first ngmax neighbors, ascending j) row-major into the wave's scratch slice (64 * ngmax slots, reused group after
 * group), the wave's stores are made visible to its own lanes, each lane folds its body's list in order and the
 * postamble result is stored. Bodies outside the subgroups take no part.
 */
template<bool AnyPbc, class T, class InTuple, class OutTuple, class Interaction, class Postamble>
__global__ __launch_bounds__(kFusedWaves * 64) void fusedIjLoopKernel(
    Box<T> box, const int32_t* __restrict__ childOffsets, const int32_t* __restrict__ internalToLeaf,
    const LocalIndex* __restrict__ layout, const T* __restrict__ centers, const T* __restrict__ sizes,
    int32_t numNodes, float ext, const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
    const T* __restrict__ h, const Record<T>* __restrict__ rec, InTuple input, OutTuple output,
    Interaction interaction, Postamble postamble, unsigned ngmax, LocalIndex first, LocalIndex last,
    GroupView groups, LocalIndex* __restrict__ scratch, uint32_t* __restrict__ counters,
    uint32_t* __restrict__ faults, const nbwalk::ChildGroup* __restrict__ childGroups)
{
    using namespace hipcstone;
    // the tree's arrays come in as separate restrict-qualified arguments: read-only memory that nothing aliases,
    // so the walk's wave-uniform node reads become scalar loads
    const nbwalk::Tree<T> tree{childOffsets, internalToLeaf, layout, centers, sizes, numNodes, ext, childGroups};
    constexpr auto inIdx = std::make_index_sequence<std::tuple_size_v<InTuple>>{};
    __shared__ nbwalk::Shared<T, kFusedWaves> sh;
    const int w             = uniform(int(threadIdx.x) / kWave);
    const int lane          = laneId();
    LocalIndex* rows        = scratch + (size_t(blockIdx.x) * kFusedWaves + size_t(w)) * kWave * ngmax;
    const uint32_t rowBytes = uint32_t(kWave) * ngmax * uint32_t(sizeof(LocalIndex));
    const uint32_t numGrp   = uint32_t((uint64_t(last) - first + kWave - 1) / kWave);
    const uint32_t xcd      = blockIdx.x % kXcds;

    for (uint32_t r = 0; r < kXcds; ++r)
    {
        const uint32_t part = (xcd + r) % kXcds;
        const uint32_t gb   = uint32_t(uint64_t(numGrp) * part / kXcds);
        const uint32_t ge   = uint32_t(uint64_t(numGrp) * (part + 1) / kXcds);
        while (true)
        {
            uint32_t g = 0;
            if (lane == 0) g = atomicAdd(&counters[part], 1u);
            g = uniform(g);
            if (g >= ge - gb) break;
            const uint64_t gFirst = uint64_t(first) + uint64_t(gb + g) * kWave;
            const LocalIndex i    = LocalIndex(gFirst) + LocalIndex(lane);
            const bool valid      = i < last && inGroups(groups, i);
            if (ballot(valid) == 0) continue;

            // the walk writes the 64 lists (first ngmax neighbors, ascending j) row-major into the wave's slice
            const uint32_t count = nbwalk::walk<T, AnyPbc, false, false, kFusedWaves>(
                x, y, z, h, box, tree, 0, gFirst, valid, ngmax, ngmax ? rows : nullptr, rowBytes, sh, w, faults);
            // the lists were stored by the lanes that tested each pair: make them visible to the lanes that fold them
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");

            // each lane folds its body's list: the self term first, then the neighbors in list order. Lanes without
            // a body fold an empty list of body `first` and store nothing.
            const LocalIndex iOwn = valid ? i : first;
            const auto iData      = loadParticleData(rec, input, iOwn, inIdx);
            const Vec3<T> pi      = std::get<1>(iData);
            bool usePbc           = false;
            if constexpr (AnyPbc) usePbc = requiresPbc(box, pi, std::get<2>(iData));
            auto result           = interaction(iData, iData, Vec3<T>{{T(0), T(0), T(0)}}, T(0));
            constexpr auto idx    = std::make_index_sequence<std::tuple_size_v<decltype(result)>>{};
            const unsigned stored = valid ? (count < ngmax ? count : ngmax) : 0u;
            // the list is read kFoldBatch neighbors at a time, the next batch's indices loaded while the current
            // batch's data is gathered, so a batch waits for one gather instead of two dependent loads per neighbor;
            // slots past the list load the body itself (a valid address) and are not folded
            const LocalIndex* myRow = rows + size_t(lane) * ngmax;
            auto slotIndex          = [&](unsigned k) { return k < stored ? myRow[k] : iOwn; };
            LocalIndex nextJ[kFoldBatch];
#pragma unroll
            for (unsigned u = 0; u < kFoldBatch; ++u)
                nextJ[u] = slotIndex(u);
            for (unsigned nb = 0; nb < stored; nb += kFoldBatch)
            {
                LocalIndex curJ[kFoldBatch];
#pragma unroll
                for (unsigned u = 0; u < kFoldBatch; ++u)
                {
                    curJ[u]  = nextJ[u];
                    nextJ[u] = slotIndex(nb + kFoldBatch + u);
                }
                decltype(loadParticleData(rec, input, iOwn, inIdx)) jData[kFoldBatch];
#pragma unroll
                for (unsigned u = 0; u < kFoldBatch; ++u)
                    jData[u] = loadParticleData(rec, input, curJ[u], inIdx);
#pragma unroll
                for (unsigned u = 0; u < kFoldBatch; ++u)
                {
                    if (nb + u < stored)
                    {
                        Vec3<T> rij;
                        T d2;
                        posDiffAndDistSq(usePbc, box, pi, std::get<1>(jData[u]), rij, d2);
                        updateResult(result, interaction(iData, jData[u], rij, d2), idx);
                    }
                }
            }
            // (every load of the slice above has completed, its value used, before the next walk writes the slice)
            if (valid)
            {
                const auto out = postamble(iData, unwrapModifiers(result, idx));
                storeParticleData(output, i, out, std::make_index_sequence<std::tuple_size_v<OutTuple>>{});
            }
        }
    }
}

inline void check(hipcstone_status st, const char* what)
{
    if (st != HIPCSTONE_OK) throw std::runtime_error(std::string(what) + ": " + hipcstone_last_error());
}

inline void checkHip(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct DeviceFree
{
    void operator()(void* p) const { (void)hipFree(p); }
};

} // namespace detail

// gpu_alwaystraverse.cuh:99-178
template<class T, class KeyType>
class GpuAlwaysTraverseNeighborhood
{
public:
    static constexpr std::size_t kScratchBytes = std::size_t(1) << 31; // the neighbor scratch's cap

    // totalBodies: the length of x, y, z, h (neighbors lie anywhere in it)
    GpuAlwaysTraverseNeighborhood(hipStream_t stream, const OctreeNsView<T, KeyType>& tree, const Box<T>& box,
                                  LocalIndex totalBodies, const GroupView& groups, const T* x, const T* y, const T* z,
                                  const T* h, unsigned ngmax)
        : stream_(stream), tree_(tree), box_(box), groups_(groups), x_(x), y_(y), z_(z), h_(h), ngmax_(ngmax),
          total_(std::max(totalBodies, groups.lastBody))
    {
        if (ngmax > (1u << 20)) throw std::runtime_error("ijloop: ngmax above 2^20");
        void* p = nullptr;
        detail::checkHip(hipMalloc(&p, hipcstone::kXcds * sizeof(uint32_t)), "ijloop: group counters");
        counters_.reset(static_cast<uint32_t*>(p));
        detail::check(hipcstone_fault_word(&faults_), "ijloop: fault word");
    }

    template<class... In, class... Out, class Interaction, class Postamble>
    void ijLoop(std::tuple<In*...> const& input, std::tuple<Out*...> const& output, Interaction&& interaction,
                Postamble&& postamble) const
    {
        run(input, output, std::forward<Interaction>(interaction), std::forward<Postamble>(postamble),
            GroupView{groups_.firstBody, groups_.lastBody, 0, nullptr, nullptr});
    }

    Statistics stats() const
    {
        return {std::size_t(groups_.lastBody - groups_.firstBody),
                scratchBytes_ + recordBytes_ + childGroupBytes_ + hipcstone::kXcds * sizeof(uint32_t)};
    }

    struct Subgroup
    {
        GpuAlwaysTraverseNeighborhood const& parent;
        GroupView groups;

        template<class... In, class... Out, class Interaction, class Postamble>
        void ijLoop(std::tuple<In*...> const& input, std::tuple<Out*...> const& output, Interaction&& interaction,
                    Postamble&& postamble) const
        {
            parent.run(input, output, std::forward<Interaction>(interaction), std::forward<Postamble>(postamble),
                       groups);
        }
    };

    // groups: device arrays groupStart/groupEnd of ascending, disjoint body ranges inside the built range
    Subgroup subgroup(GroupView const& groups) const { return {*this, groups}; }

private:
    template<class... In, class... Out, class Interaction, class Postamble>
    void run(std::tuple<In*...> const& input, std::tuple<Out*...> const& output, Interaction&& interaction,
             Postamble&& postamble, GroupView groups) const
    {
        const LocalIndex first = std::max(groups.firstBody, groups_.firstBody);
        const LocalIndex last  = std::min(groups.lastBody, groups_.lastBody);
        if (last <= first) return;
        std::tuple<const std::remove_const_t<In>*...> cin = input;
        const bool anyPbc = box_.bnd[0] == HIPCSTONE_PERIODIC || box_.bnd[1] == HIPCSTONE_PERIODIC ||
                            box_.bnd[2] == HIPCSTONE_PERIODIC;
        if (anyPbc) launch<true>(cin, output, interaction, postamble, groups, first, last);
        else launch<false>(cin, output, interaction, postamble, groups, first, last);
    }

    template<bool AnyPbc, class InTuple, class OutTuple, class Interaction, class Postamble>
    void launch(InTuple const& input, OutTuple const& output, Interaction const& interaction,
                Postamble const& postamble, GroupView groups, LocalIndex first, LocalIndex last) const
    {
        constexpr int kThreads = detail::kFusedWaves * 64;
        auto kernel = &detail::fusedIjLoopKernel<AnyPbc, T, InTuple, OutTuple, Interaction, Postamble>;
        // a persistent grid: as many workgroups as fit on the device at once (whole rounds over the 8 XCDs), no more
        // than the groups of 64 bodies need, and no more than a 2 GiB scratch holds (each wave owns a slice of
        // 64 * ngmax neighbor slots)
        int device = 0, cus = 0, perCu = 0;
        detail::checkHip(hipGetDevice(&device), "ijloop: device");
        detail::checkHip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "ijloop: CUs");
        detail::checkHip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, kThreads, 0),
                         "ijloop: occupancy");
        const std::size_t groups64  = (std::size_t(last - first) + 63) / 64;
        const std::size_t waveBytes = std::size_t(64) * ngmax_ * sizeof(LocalIndex);
        std::size_t blocks          = std::size_t(std::max(perCu, 1)) * std::size_t(std::max(cus, 1));
        blocks = std::min(blocks, (groups64 + detail::kFusedWaves - 1) / detail::kFusedWaves);
        if (waveBytes) blocks = std::min(blocks, std::max<std::size_t>(kScratchBytes / waveBytes / detail::kFusedWaves, 1));
        if (blocks > hipcstone::kXcds) blocks = blocks / hipcstone::kXcds * hipcstone::kXcds;
        const std::size_t need = blocks * detail::kFusedWaves * waveBytes;
        if (need > scratchBytes_)
        {
            scratch_.reset();
            scratchBytes_ = 0;
            void* p       = nullptr;
            detail::checkHip(hipMalloc(&p, need), "ijloop: neighbor scratch");
            scratch_.reset(static_cast<LocalIndex*>(p));
            scratchBytes_ = need;
        }
        const std::size_t recBytes = std::max<std::size_t>(total_, 1) * sizeof(detail::Record<T>);
        if (recBytes > recordBytes_)
        {
            records_.reset();
            recordBytes_ = 0;
            void* p      = nullptr;
            detail::checkHip(hipMalloc(&p, recBytes), "ijloop: record scratch");
            records_.reset(p);
            recordBytes_ = recBytes;
        }
        auto* rec = static_cast<detail::Record<T>*>(records_.get());
        detail::stageRecordsKernel<T><<<unsigned((std::size_t(total_) + 255) / 256), 256, 0, stream_>>>(
            x_, y_, z_, h_, total_, rec);
        detail::checkHip(hipGetLastError(), "ijloop: record staging");
        detail::zeroCountersKernel<<<1, 64, 0, stream_>>>(counters_.get());
        detail::checkHip(hipGetLastError(), "ijloop: group counters");
        // fp32 open boxes: the compact child blocks the walk reads (built once per neighborhood, the tree is fixed)
        const nbwalk::ChildGroup* childGroups = nullptr;
        if constexpr (std::is_same_v<T, float> && !AnyPbc)
        {
            const std::size_t count = nbwalk::childGroupCount(tree_.numNodes);
            if (count > 0 && !childGroups_)
            {
                void* p = nullptr;
                detail::checkHip(hipMalloc(&p, count * sizeof(nbwalk::ChildGroup)), "ijloop: child blocks");
                childGroups_.reset(p);
                childGroupBytes_ = count * sizeof(nbwalk::ChildGroup);
                nbwalk::childGroupsKernel<><<<unsigned((count + 255) / 256), 256, 0, stream_>>>(
                    reinterpret_cast<const float*>(tree_.centers), reinterpret_cast<const float*>(tree_.sizes),
                    tree_.childOffsets, tree_.internalToLeaf, uint32_t(count),
                    static_cast<nbwalk::ChildGroup*>(childGroups_.get()));
                detail::checkHip(hipGetLastError(), "ijloop: child blocks");
            }
            childGroups = static_cast<const nbwalk::ChildGroup*>(childGroups_.get());
        }
        kernel<<<unsigned(blocks), kThreads, 0, stream_>>>(
            box_, tree_.childOffsets, tree_.internalToLeaf, tree_.layout, reinterpret_cast<const T*>(tree_.centers),
            reinterpret_cast<const T*>(tree_.sizes), tree_.numNodes, tree_.searchExtFactor, x_, y_, z_, h_,
            rec, input, output, interaction, postamble, ngmax_, first, last, groups, scratch_.get(),
            counters_.get(), faults_, childGroups);
        detail::checkHip(hipGetLastError(), "ijloop: fused search and interaction kernel");
    }

    hipStream_t stream_;
    OctreeNsView<T, KeyType> tree_;
    Box<T> box_;
    GroupView groups_;
    const T *x_, *y_, *z_, *h_;
    unsigned ngmax_;
    LocalIndex total_;
    uint32_t* faults_ = nullptr;
    mutable std::size_t scratchBytes_ = 0;
    mutable std::unique_ptr<LocalIndex, detail::DeviceFree> scratch_;
    std::unique_ptr<uint32_t, detail::DeviceFree> counters_;
    mutable std::size_t recordBytes_ = 0;
    mutable std::unique_ptr<void, detail::DeviceFree> records_;
    mutable std::size_t childGroupBytes_ = 0;
    mutable std::unique_ptr<void, detail::DeviceFree> childGroups_;
};

// gpu_alwaystraverse.cuh:180-206
struct GpuAlwaysTraverseNeighborhoodBuilder
{
    unsigned ngmax;

    template<class T, class KeyType>
    GpuAlwaysTraverseNeighborhood<T, KeyType> build(hipStream_t stream, const OctreeNsView<T, KeyType>& tree,
                                                     const Box<T>& box, LocalIndex totalBodies,
                                                     const GroupView& groups, const T* x, const T* y, const T* z,
                                                     const T* h) const
    {
        return GpuAlwaysTraverseNeighborhood<T, KeyType>(stream, tree, box, totalBodies, groups, x, y, z, h, ngmax);
    }
};

} // namespace hipcstone::ijloop
