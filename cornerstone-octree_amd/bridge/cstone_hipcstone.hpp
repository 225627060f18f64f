// The reference-side drop-in of hipcstone for C++ callers of cornerstone (header).
//
// Include this after (or instead of) "cstone/domain/domain.hpp" in a translation unit of the reference's build. It
//   * specializes cstone::Domain<KeyType, T, execution::Gpu> for numRanks = 1: the constructor of domain.hpp:63-86,
//     sync of :169-218, syncGrav of :220-287, updateExpansionCenters of :366-372 and the accessors of :340-392
//     (startIndex, endIndex, nParticles, nParticlesWithHalos, box, layout, octreeProperties, exec), running the
//     one-rank domain of libhipcstone (hipcstone_domain_*), whose results equal the reference's CPU Domain;
//   * adds cstone::findNeighbors(execution::Gpu, ...) with the signature and contract of the CPU free function
//     findneighbors.hpp:156-177 (the reference has no GPU function with this contract).
// The free GPU functions the reference declares in its *_gpu.h headers are defined by cstone_hipcstone_bridge.cpp.
//
// Vectors passed to sync/syncGrav are device vectors with data(), size() and resize() (the reference's DeviceVector
// or any equivalent); with one rank their sizes never change, so resize() is only called on the key vector when its
// size differs from x's. Errors throw std::runtime_error, as the reference does for misuse (domain.hpp:81-85).
#pragma once

#include <array>
#include <cstring>
#include <memory>
#include <span>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "cstone/domain/domain.hpp"
#include "cstone/findneighbors.hpp"
#include "hipcstone.h"

namespace cstone
{
namespace hipcstone_bridge
{

inline void check(hipcstone_status st, const char* where)
{
    if (st != HIPCSTONE_OK) throw std::runtime_error(std::string(where) + ": " + hipcstone_last_error());
}

inline hipcstone_stream stream(execution::Gpu exec) { return static_cast<void*>(static_cast<hipStream_t>(exec)); }

template<class T>
using CBox = std::conditional_t<std::is_same_v<T, float>, hipcstone_box_f32, hipcstone_box_f64>;

// Box<T> -> the C value (the same constructor arithmetic, sfc/box.hpp:100-122)
template<class T>
CBox<T> toC(const Box<T>& b)
{
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>);
    CBox<T> out{};
    if constexpr (std::is_same_v<T, float>)
        hipcstone_make_box_f32(b.xmin(), b.xmax(), b.ymin(), b.ymax(), b.zmin(), b.zmax(), int(b.boundaryX()),
                               int(b.boundaryY()), int(b.boundaryZ()), &out);
    else
        hipcstone_make_box_f64(b.xmin(), b.xmax(), b.ymin(), b.ymax(), b.zmin(), b.zmax(), int(b.boundaryX()),
                               int(b.boundaryY()), int(b.boundaryZ()), &out);
    return out;
}

template<class V>
void* ptr(V& v)
{
    return static_cast<void*>(const_cast<std::remove_const_t<std::remove_pointer_t<decltype(v.data())>>*>(v.data()));
}

} // namespace hipcstone_bridge

/*! @brief radius neighbor search on the GPU with the CPU contract of findNeighbors (findneighbors.hpp:156-177)
 *
 * For i in [firstId, lastId): neighbors[(i - firstId) * ngmax + k] receives the ascending indices j != i with
 * |x_j - x_i|^2 < 4 h_i^2 among the leaves the particle's own traversal visits, the first ngmax of them, and
 * neighborsCount[i - firstId] the untruncated count. Every pointer (and every array of treeView) is device memory.
 */
template<class Tc, class KeyType>
void findNeighbors(execution::Gpu exec, const Tc* x, const Tc* y, const Tc* z, const Tc* h, LocalIndex firstId,
                   LocalIndex lastId, const Box<Tc>& box, const OctreeNsView<Tc, KeyType>& treeView, unsigned ngmax,
                   LocalIndex* neighbors, unsigned* neighborsCount)
{
    static_assert(sizeof(LocalIndex) == 4 && sizeof(TreeNodeIndex) == 4);
    auto cb   = hipcstone_bridge::toC(box);
    auto c    = reinterpret_cast<const Tc*>(treeView.centers);
    auto s    = reinterpret_cast<const Tc*>(treeView.sizes);
    auto strm = hipcstone_bridge::stream(exec);
    if constexpr (std::is_same_v<Tc, float>)
        hipcstone_bridge::check(hipcstone_find_neighbors_f32(strm, x, y, z, h, firstId, lastId, &cb,
                                                             treeView.childOffsets, treeView.parents,
                                                             treeView.internalToLeaf, treeView.layout, c, s,
                                                             treeView.numNodes, treeView.searchExtFactor, ngmax,
                                                             neighbors, neighborsCount),
                                "findNeighbors");
    else
        hipcstone_bridge::check(hipcstone_find_neighbors_f64(strm, x, y, z, h, firstId, lastId, &cb,
                                                             treeView.childOffsets, treeView.parents,
                                                             treeView.internalToLeaf, treeView.layout, c, s,
                                                             treeView.numNodes, treeView.searchExtFactor, ngmax,
                                                             neighbors, neighborsCount),
                                "findNeighbors");
}

/*! @brief Domain<KeyType, T, Gpu> at numRanks = 1, backed by hipcstone's one-rank domain
 *
 * Same constructor and calls as the primary template (domain/domain.hpp:36-392) for one rank; any other rank count
 * throws. The particle arrays stay in place and keep their size (no halos with one rank): startIndex() = 0,
 * endIndex() = nParticlesWithHalos() = n. octreeProperties() views the focus tree (device pointers) for
 * findNeighbors(execution::Gpu, ...).
 */
template<class KeyType, class T>
class Domain<KeyType, T, execution::Gpu>
{
    static_assert(std::is_same_v<KeyType, uint32_t> || std::is_same_v<KeyType, uint64_t>,
                  "hipcstone keys are 32 or 64 bit");
    static_assert(std::is_same_v<T, float> || std::is_same_v<T, double>, "hipcstone reals are float or double");

public:
    using RealType = T;

    Domain(execution::Gpu exec, int rank, int nRanks, unsigned bucketSize, unsigned bucketSizeFocus, float theta,
           MPI_Comm /*comm*/, const Box<T>& box = Box<T>{0, 1})
        : exec_(exec)
        , box_(box)
    {
        if (nRanks != 1 || rank != 0) throw std::runtime_error("hipcstone Domain: only numRanks == 1 is supported");
        if (bucketSize < bucketSizeFocus)
        {
            throw std::runtime_error("The bucket size of the global tree must not be smaller than the bucket size"
                                     " of the focused tree\n");
        }
        auto cb                 = hipcstone_bridge::toC(box);
        hipcstone_domain* d     = nullptr;
        hipcstone_bridge::check(hipcstone_domain_create(&d, 8 * int(sizeof(KeyType)), 8 * int(sizeof(T)), bucketSize,
                                                        bucketSizeFocus, theta, &cb),
                                "Domain");
        dom_.reset(d);
    }

    //! @brief domain.hpp:169-218 at one rank; particleProperties are reordered with x,y,z,h
    template<class KeyVec, class VectorX, class VectorH, class... Vectors1, class... Vectors2>
    void sync(KeyVec& particleKeys, VectorX& x, VectorX& y, VectorX& z, VectorH& h,
              std::tuple<Vectors1&...> particleProperties, std::tuple<Vectors2&...> /*scratchBuffers*/)
    {
        size_t n = prepare(particleKeys, x, y, z, h);
        hipcstone_bridge::check(hipcstone_domain_sync(dom_.get(), stream(), hipcstone_bridge::ptr(x),
                                                      hipcstone_bridge::ptr(y), hipcstone_bridge::ptr(z),
                                                      hipcstone_bridge::ptr(h), n,
                                                      hipcstone_bridge::ptr(particleKeys)),
                                "Domain::sync");
        gatherProperties(n, particleProperties);
        refresh();
    }

    //! @brief domain.hpp:220-287 at one rank: sync plus m reordered and the focus tree's expansion centers
    template<class KeyVec, class VectorX, class VectorH, class VectorM, class... Vectors1, class... Vectors2>
    void syncGrav(KeyVec& particleKeys, VectorX& x, VectorX& y, VectorX& z, VectorH& h, VectorM& m,
                  std::tuple<Vectors1&...> particleProperties, std::tuple<Vectors2&...> /*scratchBuffers*/)
    {
        size_t n = prepare(particleKeys, x, y, z, h);
        if (m.size() != n) throw std::runtime_error("Domain::syncGrav: m must have the size of x");
        hipcstone_bridge::check(hipcstone_domain_sync_grav(dom_.get(), stream(), hipcstone_bridge::ptr(x),
                                                           hipcstone_bridge::ptr(y), hipcstone_bridge::ptr(z),
                                                           hipcstone_bridge::ptr(h), hipcstone_bridge::ptr(m), n,
                                                           hipcstone_bridge::ptr(particleKeys)),
                                "Domain::syncGrav");
        gatherProperties(n, particleProperties);
        refresh();
    }

    //! @brief domain.hpp:366-372
    template<class VectorX, class VectorM, class VectorS1>
    void updateExpansionCenters(VectorX& x, VectorX& y, VectorX& z, VectorM& m, VectorS1& /*s1*/)
    {
        hipcstone_bridge::check(
            hipcstone_domain_update_expansion_centers(dom_.get(), stream(), hipcstone_bridge::ptr(x),
                                                      hipcstone_bridge::ptr(y), hipcstone_bridge::ptr(z),
                                                      hipcstone_bridge::ptr(m), x.size()),
            "Domain::updateExpansionCenters");
        refresh();
    }

    [[nodiscard]] LocalIndex startIndex() const { return LocalIndex(view_.start_index); }
    [[nodiscard]] LocalIndex endIndex() const { return LocalIndex(view_.end_index); }
    [[nodiscard]] LocalIndex nParticles() const { return endIndex() - startIndex(); }
    [[nodiscard]] LocalIndex nParticlesWithHalos() const { return LocalIndex(view_.num_particles); }
    [[nodiscard]] TreeNodeIndex startCell() const { return 0; }
    [[nodiscard]] TreeNodeIndex endCell() const { return view_.num_leaves; }
    //! @brief particle offsets of each focus tree leaf cell (device memory)
    std::span<const LocalIndex> layout() const
    {
        return {static_cast<const LocalIndex*>(view_.layout), size_t(view_.num_leaves) + (view_.layout ? 1 : 0)};
    }
    //! @brief the coordinate bounding box of the previous sync call
    const Box<T>& box() const { return box_; }
    KeyType assignmentStart() const { return 0; }
    //! @brief focus-tree expansion centers (4 T per node, device memory) after syncGrav / updateExpansionCenters
    const Vec4<T>* expansionCenters() const { return static_cast<const Vec4<T>*>(view_.expansion_centers); }

    OctreeNsView<T, KeyType> octreeProperties() const
    {
        return {view_.num_leaves,
                view_.num_nodes,
                static_cast<const KeyType*>(view_.prefixes),
                static_cast<const TreeNodeIndex*>(view_.child_offsets),
                static_cast<const TreeNodeIndex*>(view_.parents),
                static_cast<const TreeNodeIndex*>(view_.internal_to_leaf),
                static_cast<const TreeNodeIndex*>(view_.leaf_to_internal),
                static_cast<const TreeNodeIndex*>(view_.level_range),
                static_cast<const KeyType*>(view_.leaves),
                static_cast<const LocalIndex*>(view_.layout),
                static_cast<const Vec3<T>*>(view_.centers),
                static_cast<const Vec3<T>*>(view_.sizes)};
    }

    const execution::Gpu& exec() const { return exec_; }

private:
    struct Destroy
    {
        void operator()(hipcstone_domain* d) const { hipcstone_domain_destroy(d); }
    };

    hipcstone_stream stream() const { return hipcstone_bridge::stream(exec_); }

    template<class KeyVec, class VectorX, class VectorH>
    size_t prepare(KeyVec& keys, VectorX& x, VectorX& y, VectorX& z, VectorH& h)
    {
        static_assert(std::is_same_v<std::decay_t<decltype(*x.data())>, T>, "coordinates must be of type T");
        static_assert(std::is_same_v<std::decay_t<decltype(*h.data())>, T>, "h must be of type T");
        static_assert(sizeof(*keys.data()) == sizeof(KeyType), "keys must be of KeyType");
        size_t n = x.size();
        if (y.size() != n || z.size() != n || h.size() != n)
            throw std::runtime_error("Domain::sync: x, y, z and h must have the same size");
        if (n == 0) throw std::runtime_error("Domain::sync: no particles");
        if (keys.size() != n) keys.resize(n);
        return n;
    }

    template<class... Vectors>
    void gatherProperties(size_t n, std::tuple<Vectors&...> props)
    {
        constexpr size_t np = sizeof...(Vectors);
        if constexpr (np > 0)
        {
            std::array<void*, np> arrays{};
            std::array<size_t, np> bytes{};
            size_t i = 0;
            std::apply(
                [&](auto&... v) {
                    ((v.size() == n ? void()
                                    : throw std::runtime_error("Domain::sync: properties must have the size of x"),
                      arrays[i] = hipcstone_bridge::ptr(v), bytes[i] = sizeof(*v.data()), ++i),
                     ...);
                },
                props);
            hipcstone_bridge::check(
                hipcstone_domain_gather_properties(dom_.get(), stream(), arrays.data(), bytes.data(), int(np)),
                "Domain::sync properties");
        }
    }

    // the box of the last sync (limits as fitted by the domain, boundary types of the constructor's box)
    void refresh()
    {
        hipcstone_bridge::check(hipcstone_domain_get_view(dom_.get(), &view_), "Domain view");
        box_ = Box<T>(T(view_.box_lo[0]), T(view_.box_hi[0]), T(view_.box_lo[1]), T(view_.box_hi[1]),
                      T(view_.box_lo[2]), T(view_.box_hi[2]), box_.boundaryX(), box_.boundaryY(), box_.boundaryZ());
    }

    execution::Gpu exec_;
    Box<T> box_;
    std::unique_ptr<hipcstone_domain, Destroy> dom_;
    hipcstone_domain_view view_{};
};

} // namespace cstone
