// ijloop through the C ABI: one fixed interaction, instantiated from hipcstone_ijloop.hpp (which HIP C++ callers
// include to run their own interactions). It exercises every part of the ijloop contract — an input array, a sum,
// an integer count, min / max reductions, an odd-symmetric marker and a postamble — and is what the parity tests
// compare with the reference CPU ijloop (traversal/ijloop/cpu_alwaystraverse.hpp:90-113) running the same
// interaction.
//
//   w_ij  = (1 - d2 / (4 h_i h_i))^3                       (u * u) * u
//   rho_i = (sum over i and its neighbors j of m_j w_ij) / ((h_i h_i) h_i)
//   n_i   = 1 + number of neighbors used (at most ngmax)
//   dmin_i, dmax_i = min / max of d2 over the neighbors (the self term contributes the reductions' identities)
//   fx_i  = sum over neighbors of m_j * (x_i - x_j)          (marked odd-symmetric; a plain sum here)
#include <exception>

#include "hipcstone_ijloop.hpp"
#include "runtime.hpp"

namespace hipcstone
{

template<class T>
struct ExampleInteraction
{
    using Data = std::tuple<ijloop::LocalIndex, ijloop::Vec3<T>, T, T>;

    __device__ auto operator()(Data const& iData, Data const& jData, ijloop::Vec3<T> const& r, T d2) const
    {
        const T hi     = std::get<2>(iData);
        const T mj     = std::get<3>(jData);
        const T r2     = T(4) * hi * hi;
        const T q2     = d2 / r2;
        const T u      = T(1) - q2;
        const T w      = u * u * u;
        const bool own = std::get<0>(iData) == std::get<0>(jData);
        return std::make_tuple(mj * w, uint32_t(1),
                               ijloop::reduction::min<T>{own ? std::numeric_limits<T>::max() : d2},
                               ijloop::reduction::max<T>{own ? std::numeric_limits<T>::lowest() : d2},
                               ijloop::symmetric::odd<T>{mj * r[0]});
    }
};

template<class T>
struct ExamplePostamble
{
    using Data = std::tuple<ijloop::LocalIndex, ijloop::Vec3<T>, T, T>;

    __device__ auto operator()(Data const& iData, std::tuple<T, uint32_t, T, T, T> const& res) const
    {
        const T hi = std::get<2>(iData);
        return std::make_tuple(std::get<0>(res) / (hi * hi * hi), std::get<1>(res), std::get<2>(res),
                               std::get<3>(res), std::get<4>(res));
    }
};

template<class T>
static hipcstone_status ijloopExample(hipcstone_stream s, const int32_t* childOffsets, const int32_t* parents,
                                      const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                      const T* sizes, int32_t numNodes, const ijloop::Box<T>* box, uint32_t first,
                                      uint32_t last, const T* x, const T* y, const T* z, const T* h, const T* m,
                                      uint32_t ngmax, float ext, uint32_t totalBodies, uint32_t numGroups,
                                      const uint32_t* groupStart, const uint32_t* groupEnd, T* rho, uint32_t* count,
                                      T* dmin, T* dmax, T* fx)
{
    if (last <= first) return HIPCSTONE_OK;
    if (!childOffsets || !parents || !internalToLeaf || !layout || !centers || !sizes || !box || !x || !y || !z ||
        !h || !m || !rho || !count || !dmin || !dmax || !fx || (numGroups && (!groupStart || !groupEnd)))
        return invalid("ijloop example: null pointer");
    if (totalBodies < last) return invalid("ijloop example: total_bodies below last");
    try
    {
        ijloop::OctreeNsView<T, uint64_t> tree{};
        tree.numNodes        = numNodes;
        tree.childOffsets    = childOffsets;
        tree.parents         = parents;
        tree.internalToLeaf  = internalToLeaf;
        tree.layout          = layout;
        tree.centers         = reinterpret_cast<const ijloop::Vec3<T>*>(centers);
        tree.sizes           = reinterpret_cast<const ijloop::Vec3<T>*>(sizes);
        tree.searchExtFactor = ext;
        ijloop::GroupView groups{first, last, 0, nullptr, nullptr};
        auto nb = ijloop::GpuAlwaysTraverseNeighborhoodBuilder{ngmax}.build(stream(s), tree, *box, totalBodies,
                                                                            groups, x, y, z, h);
        if (numGroups == 0)
            nb.ijLoop(std::make_tuple(m), std::make_tuple(rho, count, dmin, dmax, fx), ExampleInteraction<T>{},
                      ExamplePostamble<T>{});
        else
            nb.subgroup(ijloop::GroupView{first, last, numGroups, groupStart, groupEnd})
                .ijLoop(std::make_tuple(m), std::make_tuple(rho, count, dmin, dmax, fx), ExampleInteraction<T>{},
                        ExamplePostamble<T>{});
    }
    catch (std::exception const& e)
    {
        setError(e.what());
        return HIPCSTONE_HIP_ERROR;
    }
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

#define HIPCSTONE_DEF_IJLOOP_EXAMPLE(RS, T)                                                                            \
    hipcstone_status hipcstone_ijloop_example_##RS(                                                                    \
        hipcstone_stream s, const int32_t* child_offsets, const int32_t* parents, const int32_t* internal_to_leaf,    \
        const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes, const hipcstone_box_##RS* box,   \
        uint32_t first, uint32_t last, const T* x, const T* y, const T* z, const T* h, const T* m, uint32_t ngmax,    \
        float ext, uint32_t total_bodies, uint32_t num_groups, const uint32_t* group_start,                          \
        const uint32_t* group_end, T* rho, uint32_t* count, T* dmin, T* dmax, T* fx)                                   \
    {                                                                                                                  \
        return ijloopExample<T>(s, child_offsets, parents, internal_to_leaf, layout, centers, sizes, num_nodes, box,  \
                                first, last, x, y, z, h, m, ngmax, ext, total_bodies, num_groups, group_start,        \
                                group_end, rho, count, dmin, dmax, fx);                                                \
    }
HIPCSTONE_DEF_IJLOOP_EXAMPLE(f32, float)
HIPCSTONE_DEF_IJLOOP_EXAMPLE(f64, double)

} // extern "C"
