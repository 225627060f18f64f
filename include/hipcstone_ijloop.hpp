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
 * The neighbor search runs in the library (libhipcstone.so); the interaction kernel is instantiated here, in the
 * caller's translation unit, for the caller's interaction and postamble (device-callable functors or lambdas).
 * Bodies are processed in chunks so that the neighbor scratch stays below 2 GiB; the scratch holds each group of 64
 * bodies' lists slot-major (hipcstone_find_neighbors_slot_major_*), so the fold's per-slot index loads coalesce, and
 * x, y, z, h of all totalBodies bodies are staged as one record per body at each ijLoop, so a neighbor's position
 * and h come in one wide load (the values are copies, the results do not change).
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

template<bool AnyPbc, class T, class... In, class... Out, class Interaction, class Postamble>
__global__ __launch_bounds__(256) void ijLoopKernel(Box<T> box, const Record<T>* __restrict__ rec,
                                                     std::tuple<const In*...> input, std::tuple<Out*...> output,
                                                     Interaction interaction, Postamble postamble, unsigned ngmax,
                                                     const LocalIndex* __restrict__ neighbors,
                                                     const LocalIndex* __restrict__ counts, LocalIndex chunkFirst,
                                                     LocalIndex chunkLast, GroupView groups)
{
    auto body = [&](LocalIndex i) {
        const auto iData = loadParticleData(rec, input, i, std::index_sequence_for<In...>{});
        const Vec3<T> pi = std::get<1>(iData);
        bool usePbc      = false;
        if constexpr (AnyPbc) usePbc = requiresPbc(box, pi, std::get<2>(iData));
        const LocalIndex row = i - chunkFirst;
        const unsigned nbs   = counts[row] < ngmax ? counts[row] : ngmax;
        // slot-major scratch: neighbor nb of row r at ((r / 64) * ngmax + nb) * 64 + r % 64
        const LocalIndex* slots = neighbors + size_t(row >> 6) * ngmax * 64 + (row & 63u);
        auto result          = interaction(iData, iData, Vec3<T>{{T(0), T(0), T(0)}}, T(0));
        constexpr auto idx   = std::make_index_sequence<std::tuple_size_v<decltype(result)>>{};
        for (unsigned nb = 0; nb < nbs; ++nb)
        {
            const LocalIndex j = slots[size_t(nb) * 64];
            const auto jData   = loadParticleData(rec, input, j, std::index_sequence_for<In...>{});
            Vec3<T> r;
            T d2;
            posDiffAndDistSq(usePbc, box, pi, std::get<1>(jData), r, d2);
            updateResult(result, interaction(iData, jData, r, d2), idx);
        }
        const auto out = postamble(iData, unwrapModifiers(result, idx));
        storeParticleData(output, i, out, std::index_sequence_for<Out...>{});
    };

    if (groups.numGroups == 0)
    {
        LocalIndex i = chunkFirst + LocalIndex(blockIdx.x) * blockDim.x + threadIdx.x;
        if (i < chunkLast) body(i);
        return;
    }
    // subgroups: block g walks group g's bodies inside this chunk
    for (LocalIndex g = blockIdx.x; g < groups.numGroups; g += gridDim.x)
    {
        LocalIndex b = groups.groupStart[g] > chunkFirst ? groups.groupStart[g] : chunkFirst;
        LocalIndex e = groups.groupEnd[g] < chunkLast ? groups.groupEnd[g] : chunkLast;
        for (LocalIndex i = b + threadIdx.x; i < e; i += blockDim.x)
            body(i);
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

inline hipcstone_status findNeighbors(hipStream_t s, const float* x, const float* y, const float* z, const float* h,
                                      LocalIndex first, LocalIndex last, const Box<float>& box,
                                      const int32_t* childOffsets, const int32_t* parents,
                                      const int32_t* internalToLeaf, const LocalIndex* layout, const void* centers,
                                      const void* sizes, int32_t numNodes, float ext, unsigned ngmax,
                                      LocalIndex* nb, LocalIndex* nc)
{
    return hipcstone_find_neighbors_slot_major_f32(reinterpret_cast<hipcstone_stream>(s), x, y, z, h, first, last,
                                                   &box,
                                        childOffsets, parents, internalToLeaf, layout,
                                        static_cast<const float*>(centers), static_cast<const float*>(sizes),
                                        numNodes, ext, ngmax, nb, nc);
}

inline hipcstone_status findNeighbors(hipStream_t s, const double* x, const double* y, const double* z,
                                      const double* h, LocalIndex first, LocalIndex last, const Box<double>& box,
                                      const int32_t* childOffsets, const int32_t* parents,
                                      const int32_t* internalToLeaf, const LocalIndex* layout, const void* centers,
                                      const void* sizes, int32_t numNodes, float ext, unsigned ngmax,
                                      LocalIndex* nb, LocalIndex* nc)
{
    return hipcstone_find_neighbors_slot_major_f64(reinterpret_cast<hipcstone_stream>(s), x, y, z, h, first, last,
                                                   &box,
                                        childOffsets, parents, internalToLeaf, layout,
                                        static_cast<const double*>(centers), static_cast<const double*>(sizes),
                                        numNodes, ext, ngmax, nb, nc);
}

} // namespace detail

// gpu_alwaystraverse.cuh:99-178
template<class T, class KeyType>
class GpuAlwaysTraverseNeighborhood
{
public:
    static constexpr std::size_t kScratchElements = std::size_t(1) << 29; // neighbor slots per chunk (2 GiB)

    // totalBodies: the length of x, y, z, h (neighbors lie anywhere in it)
    GpuAlwaysTraverseNeighborhood(hipStream_t stream, const OctreeNsView<T, KeyType>& tree, const Box<T>& box,
                                  LocalIndex totalBodies, const GroupView& groups, const T* x, const T* y, const T* z,
                                  const T* h, unsigned ngmax)
        : stream_(stream), tree_(tree), box_(box), groups_(groups), x_(x), y_(y), z_(z), h_(h), ngmax_(ngmax),
          total_(std::max(totalBodies, groups.lastBody))
    {
        LocalIndex bodies = groups.lastBody > groups.firstBody ? groups.lastBody - groups.firstBody : 0;
        std::size_t fit   = std::max<std::size_t>(kScratchElements / std::max(ngmax, 1u) / 64 * 64, 64);
        chunk_            = LocalIndex(std::min<std::size_t>((std::size_t(std::max<LocalIndex>(bodies, 1)) + 63) / 64 * 64,
                                                             fit));
        void* p           = nullptr;
        detail::checkHip(hipMalloc(&p, std::max<std::size_t>(std::size_t(chunk_) * ngmax, 1) * sizeof(LocalIndex)),
                         "ijloop: neighbor scratch");
        neighbors_.reset(static_cast<LocalIndex*>(p));
        detail::checkHip(hipMalloc(&p, std::size_t(chunk_) * sizeof(LocalIndex)), "ijloop: count scratch");
        counts_.reset(static_cast<LocalIndex*>(p));
        detail::checkHip(hipMalloc(&p, std::max<std::size_t>(total_, 1) * sizeof(detail::Record<T>)),
                         "ijloop: record scratch");
        records_.reset(static_cast<detail::Record<T>*>(p));
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
                std::size_t(chunk_) * (std::size_t(ngmax_) + 1) * sizeof(LocalIndex) +
                    std::size_t(total_) * sizeof(detail::Record<T>)};
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
        detail::stageRecordsKernel<T><<<unsigned((std::size_t(total_) + 255) / 256), 256, 0, stream_>>>(
            x_, y_, z_, h_, total_, records_.get());
        detail::checkHip(hipGetLastError(), "ijloop: record staging");
        for (LocalIndex c0 = first; c0 < last; c0 += std::min(chunk_, last - c0))
        {
            const LocalIndex c1 = c0 + std::min(chunk_, last - c0);
            detail::check(detail::findNeighbors(stream_, x_, y_, z_, h_, c0, c1, box_, tree_.childOffsets,
                                                tree_.parents, tree_.internalToLeaf, tree_.layout, tree_.centers,
                                                tree_.sizes, tree_.numNodes, tree_.searchExtFactor, ngmax_,
                                                neighbors_.get(), counts_.get()),
                          "ijloop: findNeighbors");
            unsigned blocks = groups.numGroups ? std::min<unsigned>(groups.numGroups, 65535u)
                                               : unsigned((std::size_t(c1 - c0) + 255) / 256);
            if (anyPbc)
                detail::ijLoopKernel<true, T><<<blocks, 256, 0, stream_>>>(
                    box_, records_.get(), cin, output, interaction, postamble, ngmax_, neighbors_.get(),
                    counts_.get(), c0, c1, groups);
            else
                detail::ijLoopKernel<false, T><<<blocks, 256, 0, stream_>>>(
                    box_, records_.get(), cin, output, interaction, postamble, ngmax_, neighbors_.get(),
                    counts_.get(), c0, c1, groups);
            detail::checkHip(hipGetLastError(), "ijloop: interaction kernel");
        }
    }

    hipStream_t stream_;
    OctreeNsView<T, KeyType> tree_;
    Box<T> box_;
    GroupView groups_;
    const T *x_, *y_, *z_, *h_;
    unsigned ngmax_;
    LocalIndex total_;
    LocalIndex chunk_ = 1;
    std::unique_ptr<LocalIndex, detail::DeviceFree> neighbors_, counts_;
    std::unique_ptr<detail::Record<T>, detail::DeviceFree> records_;
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
