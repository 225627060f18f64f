/*
 * ORACLE — test infrastructure, not product code.
 *
 * C-callable wrappers around the reference's own CPU functions. This file contains no reference code: it
 * includes the reference headers where they lie (-I /root/reference/include, see oracle/Makefile) and is built
 * into oracle/_ref/libcstone_ref.so, which only tests/ and bench.py's cpu_baseline leg load, to pin the
 * oracle's C restatement and to time the reference CPU path. Argument layouts match oracle.c.
 */
#include <chrono>
#include <omp.h>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

#include "cstone/domain/domaindecomp.hpp"
#include "cstone/domain/layout.hpp"
#include "cstone/findneighbors.hpp"
#include "cstone/focus/octree_focus.hpp"
#include "cstone/focus/source_center.hpp"
#include "cstone/primitives/gather.hpp"
#include "cstone/sfc/sfc.hpp"
#include "cstone/traversal/collisions.hpp"
#include "cstone/traversal/ijloop/cpu_alwaystraverse.hpp"
#include "cstone/tree/csarray.hpp"
#include "cstone/tree/octree.hpp"

using namespace cstone;

namespace
{

template<class T>
struct BoxC
{
    T lim[6];
    T len[3];
    T ilen[3];
    uint8_t bnd[3];
};

template<class T>
Box<T> toBox(const BoxC<T>* b)
{
    return Box<T>(b->lim[0], b->lim[1], b->lim[2], b->lim[3], b->lim[4], b->lim[5], BoundaryType(b->bnd[0]),
                  BoundaryType(b->bnd[1]), BoundaryType(b->bnd[2]));
}

template<class T>
const Vec3<T>* v3(const T* p)
{
    return reinterpret_cast<const Vec3<T>*>(p);
}

template<class T>
Vec3<T>* v3(T* p)
{
    return reinterpret_cast<Vec3<T>*>(p);
}

template<class K>
void keysImpl(int curve, const void* x, const void* y, const void* z, K* keys, size_t n, const void* box, int isDouble)
{
    if (isDouble)
    {
        auto b = toBox(static_cast<const BoxC<double>*>(box));
        if (curve == 1)
            computeSfcKeys((const double*)x, (const double*)y, (const double*)z, (HilbertKey<K>*)keys, n, b);
        else
            computeSfcKeys((const double*)x, (const double*)y, (const double*)z, (MortonKey<K>*)keys, n, b);
    }
    else
    {
        auto b = toBox(static_cast<const BoxC<float>*>(box));
        if (curve == 1)
            computeSfcKeys((const float*)x, (const float*)y, (const float*)z, (HilbertKey<K>*)keys, n, b);
        else
            computeSfcKeys((const float*)x, (const float*)y, (const float*)z, (MortonKey<K>*)keys, n, b);
    }
}

template<class K>
int computeOctreeImpl(const K* keys, size_t n, unsigned bucket, K* treeOut, int cap, unsigned* countsOut)
{
    auto [tree, counts] = computeOctree<K>(std::span<const K>(keys, n), bucket);
    int nn              = int(nNodes(tree));
    if (nn <= cap)
    {
        std::copy(tree.begin(), tree.end(), treeOut);
        std::copy(counts.begin(), counts.end(), countsOut);
    }
    return nn;
}

template<class K>
int updateOctreeImpl(const K* keys, size_t n, unsigned bucket, K* tree, int* numNodes, unsigned* counts, int cap)
{
    std::vector<K> t(tree, tree + *numNodes + 1);
    std::vector<unsigned> c(counts, counts + *numNodes);
    bool conv = updateOctree<K>(std::span<const K>(keys, n), bucket, t, c);
    int nn    = int(nNodes(t));
    if (nn <= cap)
    {
        std::copy(t.begin(), t.end(), tree);
        std::copy(c.begin(), c.end(), counts);
    }
    *numNodes = nn;
    return conv ? 1 : 0;
}

template<class K>
void buildOctreeImpl(const K* leaves, int numLeaves, K* prefixes, int32_t* childOffsets, int32_t* parents,
                     int32_t* levelRange, int32_t* internalToLeaf, int32_t* leafToInternal)
{
    int numInternal = (numLeaves - 1) / 7;
    buildOctreeCpu(leaves, numLeaves, numInternal, prefixes, childOffsets, parents, levelRange, internalToLeaf,
                   leafToInternal);
}

template<class T>
void findNeighborsImpl(const T* x, const T* y, const T* z, const T* h, uint32_t first, uint32_t last,
                       const BoxC<T>* box, const int32_t* childOffsets, const int32_t* parents,
                       const int32_t* internalToLeaf, const uint32_t* layout, const T* centers, const T* sizes,
                       float ext, unsigned ngmax, uint32_t* nb, uint32_t* nc)
{
    OctreeNsView<T, uint64_t> v{};
    v.childOffsets    = childOffsets;
    v.parents         = parents;
    v.internalToLeaf  = internalToLeaf;
    v.layout          = layout;
    v.centers         = v3(centers);
    v.sizes           = v3(sizes);
    v.searchExtFactor = ext;
    findNeighbors(x, y, z, h, first, last, toBox(box), v, ngmax, nb, nc);
}

// The reference's CPU ijloop (CpuAlwaysTraverseNeighborhoodBuilder) running the example interaction of
// hipcstone_ijloop_example_*: w = (1 - d2 / r2)^3 with r2 = radiusSq-style (4 h_i) h_i, sums, count, min / max
// reductions, an odd-symmetric marker, and a postamble dividing by (h_i h_i) h_i.
template<class T>
void ijloopExampleImpl(const T* x, const T* y, const T* z, const T* h, const T* m, uint32_t first, uint32_t last,
                       const BoxC<T>* box, const int32_t* childOffsets, const int32_t* parents,
                       const int32_t* internalToLeaf, const uint32_t* layout, const T* centers, const T* sizes,
                       float ext, unsigned ngmax, T* rho, uint32_t* count, T* dmin, T* dmax, T* fx)
{
    OctreeNsView<T, uint64_t> v{};
    v.childOffsets    = childOffsets;
    v.parents         = parents;
    v.internalToLeaf  = internalToLeaf;
    v.layout          = layout;
    v.centers         = v3(centers);
    v.sizes           = v3(sizes);
    v.searchExtFactor = ext;
    GroupView groups{first, last, 0, nullptr, nullptr};
    auto interaction = [](auto const& iData, auto const& jData, Vec3<T> const& r, T d2)
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
    };
    auto postamble = [](auto const& iData, std::tuple<T, uint32_t, T, T, T> const& res)
    {
        const T hi = std::get<2>(iData);
        return std::make_tuple(std::get<0>(res) / (hi * hi * hi), std::get<1>(res), std::get<2>(res),
                               std::get<3>(res), std::get<4>(res));
    };
    ijloop::CpuAlwaysTraverseNeighborhoodBuilder{ngmax}
        .build(execution::Cpu{}, v, toBox(box), last, groups, x, y, z, h)
        .ijLoop(std::make_tuple(m), std::make_tuple(rho, count, dmin, dmax, fx), interaction, postamble);
}

template<class K, class T>
void findHalosImpl(const K* prefixes, const int32_t* childOffsets, const int32_t* parents, const T* nodeCenters,
                   const T* nodeSizes, const K* leaves, const T* searchCenters, const T* searchSizes,
                   const BoxC<T>* box, int32_t firstNode, int32_t lastNode, uint8_t* flags)
{
    findHalos(prefixes, childOffsets, parents, v3(nodeCenters), v3(nodeSizes), leaves, v3(searchCenters),
              v3(searchSizes), toBox(box), firstNode, lastNode, flags);
}

template<class K, class T>
void centersImpl(int /*curve: the reference decodes node keys as Hilbert keys (sfcKey -> SfcKind)*/, const K* prefixes,
                 size_t n, T* centers, T* sizes, const BoxC<T>* box)
{
    nodeFpCenters<K>(std::span<const K>(prefixes, n), v3(centers), v3(sizes), toBox(box));
}

template<class T>
void bboxImpl(const T* x, const T* y, const T* z, const T* h, const uint32_t* layout, int32_t first, int32_t last,
              T scale, T* centers, T* sizes)
{
#pragma omp parallel for
    for (int32_t leaf = first; leaf < last; ++leaf)
    {
        auto [c, s] = computeBoundingBox(x, y, z, h, layout[leaf], layout[leaf + 1], scale, v3(centers)[leaf]);
        v3(centers)[leaf] = c;
        v3(sizes)[leaf]   = s;
    }
}

double nowMs()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

/*
 * One step of the plain pipeline as the reference's CPU path runs it (the sequence of
 * test/unit/neighbors/findneighbors.cpp:43-93 and test/performance/octree.cpp:29-100):
 * keys, stable sort with iota, gather x,y,z,h, updateOctree (warm start from tree/counts, or computeOctree when
 * numNodes == 0), updateInternalTree, layout scan, nodeFpCenters, then the reference's per-particle findNeighbors
 * (findneighbors.hpp:77-154) for every nbStride-th particle, in an OpenMP loop like the batch version (:156-177).
 * Arrays x,y,z,h are reordered in place. Stage times (ms) go to times[0..6].
 */
template<class K, class T>
int pipelineImpl(int curve, T* x, T* y, T* z, T* h, size_t n, const BoxC<T>* boxc, unsigned bucket, K* tree,
                 int* numNodes, unsigned* counts, int cap, unsigned ngmax, uint32_t nbStride, uint32_t* nb,
                 uint32_t* nc, double* times)
{
    Box<T> box = toBox(boxc);
    std::vector<K> keys(n);
    double t0 = nowMs();
    if (curve == 1) computeSfcKeys(x, y, z, (HilbertKey<K>*)keys.data(), n, box);
    else computeSfcKeys(x, y, z, (MortonKey<K>*)keys.data(), n, box);
    double t1 = nowMs();
    std::vector<LocalIndex> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    sort_by_key(keys.begin(), keys.end(), perm.begin());
    std::vector<T> tmp(n);
    for (T* a : {x, y, z, h})
    {
        gather<LocalIndex>(std::span<const LocalIndex>(perm.data(), n), a, tmp.data());
        std::copy(tmp.begin(), tmp.end(), a);
    }
    double t2 = nowMs();
    std::vector<K> t;
    std::vector<unsigned> c;
    if (*numNodes == 0)
    {
        auto [tt, cc] = computeOctree<K>(std::span<const K>(keys.data(), n), bucket);
        t             = std::move(tt);
        c             = std::move(cc);
    }
    else
    {
        t.assign(tree, tree + *numNodes + 1);
        c.assign(counts, counts + *numNodes);
        updateOctree<K>(std::span<const K>(keys.data(), n), bucket, t, c);
    }
    double t3 = nowMs();
    OctreeData<K, execution::Cpu> octree;
    octree.resize(nNodes(t));
    updateInternalTree<K>(t, octree.data());
    double t4 = nowMs();
    std::vector<LocalIndex> layout(nNodes(t) + 1);
    layout[0] = 0;
    std::inclusive_scan(c.begin(), c.end(), layout.begin() + 1);
    std::vector<Vec3<T>> centers(octree.numNodes), sizes(octree.numNodes);
    nodeFpCenters<K>(std::span<const K>(octree.prefixes.data(), octree.numNodes), centers.data(), sizes.data(), box);
    double t5 = nowMs();
    OctreeNsView<T, K> v{};
    v.numLeafNodes   = octree.numLeafNodes;
    v.numNodes       = octree.numNodes;
    v.prefixes       = octree.prefixes.data();
    v.childOffsets   = octree.childOffsets.data();
    v.parents        = octree.parents.data();
    v.internalToLeaf = octree.internalToLeaf.data();
    v.leafToInternal = octree.leafToInternal.data();
    v.levelRange     = octree.levelRange.data();
    v.leaves         = t.data();
    v.layout         = layout.data();
    v.centers        = centers.data();
    v.sizes          = sizes.data();
    int64_t numTargets = (int64_t(n) + nbStride - 1) / nbStride;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < numTargets; ++k)
    {
        LocalIndex i = LocalIndex(k * nbStride);
        nc[k]        = findNeighbors(i, x, y, z, h, v, box, ngmax, nb + size_t(k) * ngmax);
    }
    double t6 = nowMs();
    times[0]  = t1 - t0;
    times[1]  = t2 - t1;
    times[2]  = t3 - t2;
    times[3]  = t4 - t3;
    times[4]  = t5 - t4;
    times[5]  = t6 - t5;
    times[6]  = t6 - t0;
    int nn    = int(nNodes(t));
    if (nn <= cap)
    {
        std::copy(t.begin(), t.end(), tree);
        std::copy(c.begin(), c.end(), counts);
    }
    *numNodes = nn;
    return nn;
}

/*
 * The rank-share chain through the reference's CPU functions, timed per stage: keys, stable sort + gather,
 * computeOctree, updateInternalTree, nodeFpCenters + layout, makeSfcAssignment, discoverHalos' CPU branch
 * (octree_focus_mpi.hpp:553-566: search boxes seeded with leaf centers, computeBoundingBox, findHalos),
 * computeNodeLayout, and the halo gather as extractMarkedElements ranges copied into local arrays.
 * times[0..7]: keys, sort+gather, tree, link, centers+layout, assignment, halos, layout+gather. Returns the number of
 * local particles.
 */
template<class K, class T>
size_t haloChainImpl(int curve, T* x, T* y, T* z, T* h, size_t n, const BoxC<T>* boxc, unsigned bucket, int numRanks,
                     int rank, float searchExt, double* times)
{
    Box<T> box = toBox(boxc);
    std::vector<K> keys(n);
    double t0 = nowMs();
    if (curve == 1) computeSfcKeys(x, y, z, (HilbertKey<K>*)keys.data(), n, box);
    else computeSfcKeys(x, y, z, (MortonKey<K>*)keys.data(), n, box);
    double t1 = nowMs();
    std::vector<LocalIndex> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    sort_by_key(keys.begin(), keys.end(), perm.begin());
    std::vector<T> tmp(n);
    for (T* a : {x, y, z, h})
    {
        gather<LocalIndex>(std::span<const LocalIndex>(perm.data(), n), a, tmp.data());
        std::copy(tmp.begin(), tmp.end(), a);
    }
    double t2 = nowMs();
    auto [tree, counts] = computeOctree<K>(std::span<const K>(keys.data(), n), bucket);
    double t3 = nowMs();
    OctreeData<K, execution::Cpu> octree;
    octree.resize(nNodes(tree));
    updateInternalTree<K>(tree, octree.data());
    double t4 = nowMs();
    TreeNodeIndex L = nNodes(tree);
    std::vector<LocalIndex> layout(L + 1);
    layout[0] = 0;
    std::inclusive_scan(counts.begin(), counts.end(), layout.begin() + 1);
    std::vector<Vec3<T>> centers(octree.numNodes), sizes(octree.numNodes);
    nodeFpCenters<K>(std::span<const K>(octree.prefixes.data(), octree.numNodes), centers.data(), sizes.data(), box);
    double t5 = nowMs();
    auto assignment = makeSfcAssignment(numRanks, counts, tree.data());
    TreeNodeIndex first = assignment.treeOffsetsConst()[rank], last = assignment.treeOffsetsConst()[rank + 1];
    double t6 = nowMs();
    std::vector<Vec3<T>> sc(L), ss(L);
    auto l2i = leafToInternal(octree);
    gather<TreeNodeIndex>(l2i, centers.data(), sc.data());
#pragma omp parallel for schedule(static)
    for (TreeNodeIndex i = first; i < last; ++i)
        std::tie(sc[i], ss[i]) = computeBoundingBox(x, y, z, h, layout[i], layout[i + 1], T(2 * searchExt), sc[i]);
    std::vector<uint8_t> flags(octree.numNodes, 0);
    findHalos(octree.prefixes.data(), octree.childOffsets.data(), octree.parents.data(), centers.data(),
              sizes.data(), tree.data(), sc.data(), ss.data(), box, first, last, flags.data());
    double t7 = nowMs();
    std::vector<LocalIndex> local(L + 1);
    computeNodeLayout(execution::Cpu{}, std::span<const unsigned>(counts.data(), L),
                      std::span<const uint8_t>(flags.data(), flags.size()), l2i, TreeIndexPair(first, last),
                      std::span<LocalIndex>(local.data(), local.size()));
    auto ranges = extractMarkedElements<LocalIndex>(std::span<const LocalIndex>(layout.data(), layout.size()),
                                                   std::span<const LocalIndex>(local.data(), local.size()), 0, L);
    size_t numLocal = local[L];
    std::vector<size_t> dstOff(ranges.size() / 2 + 1, 0);
    for (size_t r = 0; r + 1 < ranges.size(); r += 2)
        dstOff[r / 2 + 1] = dstOff[r / 2] + (ranges[r + 1] - ranges[r]);
    std::vector<T> outs[4];
    T* srcs[4] = {x, y, z, h};
    for (int a = 0; a < 4; ++a)
    {
        outs[a].resize(numLocal);
#pragma omp parallel for schedule(dynamic, 64)
        for (size_t r = 0; r < ranges.size() / 2; ++r)
            std::copy(srcs[a] + ranges[2 * r], srcs[a] + ranges[2 * r + 1], outs[a].data() + dstOff[r]);
    }
    double t8 = nowMs();
    double tt[8] = {t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4, t6 - t5, t7 - t6, t8 - t7};
    std::copy(tt, tt + 8, times);
    return numLocal;
}


template<class T>
void leafMassCentersImpl(const T* x, const T* y, const T* z, const T* m, const int32_t* leafToInternal,
                         int32_t numLeaves, const uint32_t* layout, T* centers)
{
    size_t n = layout[numLeaves];
    computeLeafMassCenter<T, T, T>(std::span<const T>(x, n), std::span<const T>(y, n), std::span<const T>(z, n),
                                   std::span<const T>(m, n), std::span<const TreeNodeIndex>(leafToInternal, numLeaves),
                                   layout, reinterpret_cast<SourceCenterType<T>*>(centers));
}

template<class T>
void upsweepCentersImpl(const int32_t* levelRange, int numLevels, const int32_t* childOffsets, T* centers)
{
    upsweep(std::span<const TreeNodeIndex>(levelRange, size_t(numLevels) + 2), childOffsets,
            reinterpret_cast<SourceCenterType<T>*>(centers), CombineSourceCenter<T>{});
}

template<class K, class T>
void setMacImpl(const K* prefixes, size_t n, T* centers, float invTheta, const BoxC<T>* boxc)
{
    setMac<T, K>(std::span<const K>(prefixes, n), std::span<SourceCenterType<T>>(
                                                       reinterpret_cast<SourceCenterType<T>*>(centers), n),
                 invTheta, toBox(boxc));
}

template<class K, class T>
void geoMacSpheresImpl(const K* prefixes, size_t n, T* centers, float invTheta, const BoxC<T>* boxc)
{
    geoMacSpheres<K, T>(std::span<const K>(prefixes, n), reinterpret_cast<SourceCenterType<T>*>(centers), invTheta,
                        toBox(boxc));
}

} // namespace

extern "C"
{

#define REF_KEYS(K, KS)                                                                                                \
    void ref_compute_sfc_keys_##KS(int curve, const void* x, const void* y, const void* z, K* keys, size_t n,         \
                                   const void* box, int isDouble)                                                      \
    {                                                                                                                  \
        keysImpl<K>(curve, x, y, z, keys, n, box, isDouble);                                                           \
    }                                                                                                                  \
    void ref_sort_by_key_##KS(K* keys, uint32_t* values, size_t n) { sort_by_key(keys, keys + n, values); }            \
    void ref_compute_node_counts_##KS(const K* tree, unsigned* counts, int nNodes, const K* keys, size_t nKeys,       \
                                      unsigned maxCount, int useGuess)                                                 \
    {                                                                                                                  \
        computeNodeCounts(tree, counts, nNodes, std::span<const K>(keys, nKeys), maxCount, useGuess != 0);           \
    }                                                                                                                  \
    int ref_rebalance_decision_##KS(const K* tree, const unsigned* counts, int nNodes, unsigned bucket, int32_t* ops)  \
    {                                                                                                                  \
        return rebalanceDecision(tree, counts, nNodes, bucket, ops) ? 1 : 0;                                           \
    }                                                                                                                  \
    int ref_compute_octree_##KS(const K* keys, size_t n, unsigned bucket, K* tree, int cap, unsigned* counts)         \
    {                                                                                                                  \
        return computeOctreeImpl<K>(keys, n, bucket, tree, cap, counts);                                              \
    }                                                                                                                  \
    int ref_update_octree_##KS(const K* keys, size_t n, unsigned bucket, K* tree, int* numNodes, unsigned* counts,   \
                               int cap)                                                                                \
    {                                                                                                                  \
        return updateOctreeImpl<K>(keys, n, bucket, tree, numNodes, counts, cap);                                     \
    }                                                                                                                  \
    void ref_build_octree_##KS(const K* leaves, int numLeaves, K* prefixes, int32_t* childOffsets, int32_t* parents,  \
                               int32_t* levelRange, int32_t* internalToLeaf, int32_t* leafToInternal)                  \
    {                                                                                                                  \
        buildOctreeImpl<K>(leaves, numLeaves, prefixes, childOffsets, parents, levelRange, internalToLeaf,            \
                           leafToInternal);                                                                            \
    }

    REF_KEYS(uint32_t, u32)
    REF_KEYS(uint64_t, u64)

#define REF_DOMAIN(K, KS)                                                                                              \
    void ref_make_sfc_assignment_##KS(const K* tree, const unsigned* counts, int numLeaves, int numRanks,             \
                                      int32_t* treeOffsets, K* boundaries, uint32_t* rankCounts)                       \
    {                                                                                                                  \
        std::vector<unsigned> c(counts, counts + numLeaves);                                                           \
        auto a = makeSfcAssignment(numRanks, c, tree);                                                                 \
        for (int r = 0; r <= numRanks; ++r)                                                                            \
        {                                                                                                              \
            treeOffsets[r] = a.treeOffsetsConst()[r];                                                                  \
            boundaries[r]  = a[r];                                                                                     \
        }                                                                                                              \
        for (int r = 0; r < numRanks; ++r)                                                                             \
            rankCounts[r] = a.totalCount(r);                                                                           \
    }                                                                                                                  \
    size_t ref_extract_marked_elements_##KS(const K* source, size_t sourceSize, const uint32_t* flags,                \
                                            size_t flagsSize, int32_t first, int32_t second, K* pairs,                 \
                                            size_t maxPairs)                                                           \
    {                                                                                                                  \
        auto v = extractMarkedElements<K>(std::span<const K>(source, sourceSize),                                      \
                                          std::span<const LocalIndex>(flags, flagsSize), first, second);               \
        for (size_t k = 0; k < v.size() && k < 2 * maxPairs; ++k)                                                      \
            pairs[k] = v[k];                                                                                           \
        return v.size() / 2;                                                                                           \
    }

    REF_DOMAIN(uint32_t, u32)
    REF_DOMAIN(uint64_t, u64)

#define REF_FOCUS(K, KS)                                                                                               \
    int ref_update_focus_##KS(K* leavesIO, int* numLeaves, int capacity, unsigned bucket, K focusStart, K focusEnd,    \
                              const K* mandatory, int numMandatory, const unsigned* counts, const uint8_t* macs)       \
    {                                                                                                                  \
        std::vector<K> leaves(leavesIO, leavesIO + *numLeaves + 1);                                                    \
        OctreeData<K, execution::Cpu> tree;                                                                            \
        tree.resize(nNodes(leaves));                                                                                   \
        updateInternalTree<K>(leaves, tree.data());                                                                    \
        size_t N = size_t(tree.numNodes);                                                                              \
        std::vector<uint8_t> zeroMacs(N, 0);                                                                           \
        bool conv = CombinedUpdate<K>::updateFocus(tree, leaves, bucket, focusStart, focusEnd,                         \
                                                   std::span<const K>(mandatory, numMandatory),                        \
                                                   std::span<const unsigned>(counts, N),                               \
                                                   std::span<const uint8_t>(macs ? macs : zeroMacs.data(), N));        \
        int nl = int(leaves.size()) - 1;                                                                               \
        if (nl <= capacity) std::copy(leaves.begin(), leaves.end(), leavesIO);                                         \
        *numLeaves = nl;                                                                                               \
        return conv ? 1 : 0;                                                                                           \
    }                                                                                                                  \
    void ref_focus_counts_##KS(const K* leavesIn, int L, const K* keys, size_t n, unsigned* leafCounts,               \
                               unsigned* nodeCounts)                                                                   \
    {                                                                                                                  \
        std::vector<K> leaves(leavesIn, leavesIn + L + 1);                                                             \
        OctreeData<K, execution::Cpu> tree;                                                                            \
        tree.resize(nNodes(leaves));                                                                                   \
        updateInternalTree<K>(leaves, tree.data());                                                                    \
        std::fill(leafCounts, leafCounts + L, 0u);                                                                     \
        computeNodeCounts<K>(leaves.data(), leafCounts, L, std::span<const K>(keys, n),                               \
                             std::numeric_limits<unsigned>::max(), true);                                              \
        std::fill(nodeCounts, nodeCounts + tree.numNodes, 0u);                                                         \
        scatter(leafToInternal(tree), leafCounts, nodeCounts);                                                         \
        upsweep(tree.levelRange, tree.childOffsets.data(), nodeCounts, NodeCount<unsigned>{});                         \
    }                                                                                                                  \
    int ref_compute_spanning_tree_##KS(const K* keys, int numKeys, K* out, int capacity)                              \
    {                                                                                                                  \
        auto t = computeSpanningTree<K>(std::span<const K>(keys, numKeys));                                            \
        if (int(t.size()) <= capacity) std::copy(t.begin(), t.end(), out);                                             \
        return int(t.size());                                                                                          \
    }                                                                                                                  \
    /* domain/assignment.hpp:115-123 with tree/update_mpi.hpp:64-98 at one rank (all-reduce = identity), built from \
       the reference's rebalanceDecision / rebalanceTree / computeNodeCounts */                                        \
    int ref_domain_global_update_##KS(const K* keys, size_t n, unsigned bucket, K* treeIO, int* numNodes,              \
                                      unsigned* countsIO, int capacity, int firstCall)                                 \
    {                                                                                                                  \
        std::vector<K> leaves(treeIO, treeIO + *numNodes + 1);                                                         \
        std::vector<unsigned> counts(countsIO, countsIO + *numNodes);                                                  \
        std::span<const K> keyView(keys, n);                                                                           \
        auto step = [&]() -> unsigned {                                                                                \
            std::vector<TreeNodeIndex> ops(nNodes(leaves) + 1);                                                        \
            bool conv = rebalanceDecision(leaves.data(), counts.data(), nNodes(leaves), bucket, ops.data());           \
            std::vector<K> newLeaves;                                                                                  \
            rebalanceTree(leaves, newLeaves, ops.data());                                                              \
            swap(leaves, newLeaves);                                                                                   \
            counts.resize(nNodes(leaves));                                                                             \
            computeNodeCounts(leaves.data(), counts.data(), nNodes(leaves), keyView,                                   \
                              std::numeric_limits<unsigned>::max(), true);                                             \
            if (conv) return 0u;                                                                                       \
            return *std::max_element(counts.begin(), counts.end());                                                    \
        };                                                                                                             \
        int steps  = 1;                                                                                                \
        unsigned m = step();                                                                                           \
        if (firstCall || m >= 8 * bucket)                                                                              \
        {                                                                                                              \
            ++steps;                                                                                                   \
            while (step() > bucket)                                                                                    \
                ++steps;                                                                                               \
        }                                                                                                              \
        int nn = nNodes(leaves);                                                                                       \
        if (nn <= capacity)                                                                                            \
        {                                                                                                              \
            std::copy(leaves.begin(), leaves.end(), treeIO);                                                           \
            std::copy(counts.begin(), counts.end(), countsIO);                                                         \
        }                                                                                                              \
        *numNodes = nn;                                                                                                \
        return steps;                                                                                                  \
    }

    REF_FOCUS(uint32_t, u32)
    REF_FOCUS(uint64_t, u64)

    void ref_limit_box_shrinking_f64(const double* fitting, const double* previous, const int* bnd, double* out)
    {
        auto b = [&](const double* l) {
            return Box<double>(l[0], l[1], l[2], l[3], l[4], l[5], BoundaryType(bnd[0]), BoundaryType(bnd[1]),
                               BoundaryType(bnd[2]));
        };
        auto r = limitBoxShrinking(b(fitting), b(previous));
        double v[6] = {r.xmin(), r.xmax(), r.ymin(), r.ymax(), r.zmin(), r.zmax()};
        std::copy(v, v + 6, out);
    }
    void ref_limit_box_shrinking_f32(const float* fitting, const float* previous, const int* bnd, float* out)
    {
        auto b = [&](const float* l) {
            return Box<float>(l[0], l[1], l[2], l[3], l[4], l[5], BoundaryType(bnd[0]), BoundaryType(bnd[1]),
                              BoundaryType(bnd[2]));
        };
        auto r = limitBoxShrinking(b(fitting), b(previous));
        float v[6] = {r.xmin(), r.xmax(), r.ymin(), r.ymax(), r.zmin(), r.zmax()};
        std::copy(v, v + 6, out);
    }

    void ref_compute_node_layout(const unsigned* counts, const uint8_t* flags, const int32_t* leafToInternal,
                                 int numLeaves, int numNodes, int first, int last, uint32_t* layout)
    {
        computeNodeLayout(execution::Cpu{}, std::span<const unsigned>(counts, numLeaves),
                          std::span<const uint8_t>(flags, numNodes),
                          std::span<const TreeNodeIndex>(leafToInternal, numLeaves), TreeIndexPair(first, last),
                          std::span<LocalIndex>(layout, size_t(numLeaves) + 1));
    }

#define REF_KR(K, KS, T, TS)                                                                                           \
    void ref_node_fp_centers_##KS##_##TS(int curve, const K* prefixes, size_t n, T* centers, T* sizes,               \
                                         const BoxC<T>* box)                                                           \
    {                                                                                                                  \
        centersImpl<K, T>(curve, prefixes, n, centers, sizes, box);                                                    \
    }                                                                                                                  \
    void ref_find_halos_##KS##_##TS(const K* prefixes, const int32_t* childOffsets, const int32_t* parents,           \
                                    const T* nodeCenters, const T* nodeSizes, const K* leaves, const T* searchCenters,\
                                    const T* searchSizes, const BoxC<T>* box, int32_t firstNode, int32_t lastNode,     \
                                    uint8_t* flags)                                                                    \
    {                                                                                                                  \
        findHalosImpl<K, T>(prefixes, childOffsets, parents, nodeCenters, nodeSizes, leaves, searchCenters,           \
                            searchSizes, box, firstNode, lastNode, flags);                                             \
    }                                                                                                                  \
    size_t ref_halo_chain_##KS##_##TS(int curve, T* x, T* y, T* z, T* h, size_t n, const BoxC<T>* box,             \
                                      unsigned bucket, int numRanks, int rank, float searchExt, double* times)          \
    {                                                                                                                  \
        return haloChainImpl<K, T>(curve, x, y, z, h, n, box, bucket, numRanks, rank, searchExt, times);              \
    }                                                                                                                  \
    int ref_pipeline_##KS##_##TS(int curve, T* x, T* y, T* z, T* h, size_t n, const BoxC<T>* box, unsigned bucket,   \
                                 K* tree, int* numNodes, unsigned* counts, int cap, unsigned ngmax,                    \
                                 uint32_t nbStride, uint32_t* nb, uint32_t* nc, double* times)                         \
    {                                                                                                                  \
        return pipelineImpl<K, T>(curve, x, y, z, h, n, box, bucket, tree, numNodes, counts, cap, ngmax,              \
                                  nbStride, nb, nc, times);                                                            \
    }                                                                                                                  \
    void ref_set_mac_##KS##_##TS(const K* prefixes, size_t n, T* centers, float invTheta, const BoxC<T>* box)     \
    {                                                                                                                  \
        setMacImpl<K, T>(prefixes, n, centers, invTheta, box);                                                         \
    }                                                                                                                  \
    void ref_geo_mac_spheres_##KS##_##TS(const K* prefixes, size_t n, T* centers, float invTheta,                     \
                                         const BoxC<T>* box)                                                           \
    {                                                                                                                  \
        geoMacSpheresImpl<K, T>(prefixes, n, centers, invTheta, box);                                                  \
    }

    REF_KR(uint32_t, u32, float, f32)
    REF_KR(uint32_t, u32, double, f64)
    REF_KR(uint64_t, u64, float, f32)
    REF_KR(uint64_t, u64, double, f64)

#define REF_T(T, TS)                                                                                                   \
    void ref_find_neighbors_##TS(const T* x, const T* y, const T* z, const T* h, uint32_t first, uint32_t last,       \
                                 const BoxC<T>* box, const int32_t* childOffsets, const int32_t* parents,              \
                                 const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,              \
                                 const T* sizes, float ext, unsigned ngmax, uint32_t* nb, uint32_t* nc)                \
    {                                                                                                                  \
        findNeighborsImpl<T>(x, y, z, h, first, last, box, childOffsets, parents, internalToLeaf, layout, centers,    \
                             sizes, ext, ngmax, nb, nc);                                                               \
    }                                                                                                                  \
    void ref_compute_bounding_box_##TS(const T* x, const T* y, const T* z, const T* h, const uint32_t* layout,        \
                                       int32_t first, int32_t last, T scale, T* centers, T* sizes)                     \
    {                                                                                                                  \
        bboxImpl<T>(x, y, z, h, layout, first, last, scale, centers, sizes);                                           \
    }                                                                                                                  \
    void ref_ijloop_example_##TS(const T* x, const T* y, const T* z, const T* h, const T* m, uint32_t first,         \
                                 uint32_t last, const BoxC<T>* box, const int32_t* childOffsets,                      \
                                 const int32_t* parents, const int32_t* internalToLeaf, const uint32_t* layout,        \
                                 const T* centers, const T* sizes, float ext, unsigned ngmax, T* rho,                  \
                                 uint32_t* count, T* dmin, T* dmax, T* fx)                                             \
    {                                                                                                                  \
        ijloopExampleImpl<T>(x, y, z, h, m, first, last, box, childOffsets, parents, internalToLeaf, layout, centers,  \
                             sizes, ext, ngmax, rho, count, dmin, dmax, fx);                                           \
    }                                                                                                                  \
    void ref_leaf_mass_centers_##TS(const T* x, const T* y, const T* z, const T* m, const int32_t* leafToInternal,    \
                                    int32_t numLeaves, const uint32_t* layout, T* centers)                             \
    {                                                                                                                  \
        leafMassCentersImpl<T>(x, y, z, m, leafToInternal, numLeaves, layout, centers);                               \
    }                                                                                                                  \
    void ref_upsweep_centers_##TS(const int32_t* levelRange, int numLevels, const int32_t* childOffsets, T* centers) \
    {                                                                                                                  \
        upsweepCentersImpl<T>(levelRange, numLevels, childOffsets, centers);                                           \
    }

    REF_T(float, f32)
    REF_T(double, f64)

    int ref_omp_threads()
    {
        int n = 1;
#pragma omp parallel
        {
#pragma omp single
            n = omp_get_num_threads();
        }
        return n;
    }
}
