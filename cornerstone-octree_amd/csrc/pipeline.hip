// Host orchestration of the plain single-domain pipeline (C++ above the kernels):
//   keys -> stable sort + gather x,y,z,h (later calls: through the buckets of the previous leaf array) -> updateOctree (computeOctree on the first call) -> link ->
//   geometric centers -> leaf layout -> findNeighbors,
// and of one rank's share of the synced tree: makeSfcAssignment -> discoverHalos -> computeNodeLayout -> halo
// gather (domain/domaindecomp.hpp:103-118, focus/octree_focus_mpi.hpp:511-582, halos/gather_halos_gpu.cu:28-57).
//
// This is the sequence of the reference's CPU pipeline (test/unit/neighbors/findneighbors.cpp:43-93,
// test/performance/octree.cpp:29-100) and of one Domain::sync at numRanks=1 restricted to the global tree
// (domain/domain.hpp:169-218, assignment.hpp:92-144): the first call converges the tree from the root
// (computeOctree, tree/csarray.hpp:429-440), later calls apply one warm-start updateOctree step to the previous
// leaves and counts (csarray.hpp:408-426; performance/octree.cpp:43-44).
//
// All buffers are allocated at creation or when a capacity is exceeded (outside steady-state steps). A step
// synchronizes the stream once per updateOctree call (the new leaf count decides launch sizes), once to read
// the level range and, on later calls, once to read the largest bucket of the sort.
#include <algorithm>
#include <memory>
#include <vector>

#include "common.hpp"
#include "runtime.hpp"
#include "host_util.hpp"
#include "scan.hpp"
#include "bucket_sort.hpp"


namespace hipcstone
{

struct PipelineBase
{
    StageTimer<HIPCSTONE_PIPELINE_STAGES> timer;
    virtual ~PipelineBase() = default;
    virtual hipcstone_status sync(hipcstone_stream s, const void* x, const void* y, const void* z, const void* h,
                                  size_t n, const void* box, int fullBuild, uint32_t* nb, uint32_t* nc) = 0;
    virtual hipcstone_status halos(hipcstone_stream s, int32_t numRanks, int32_t rank, float ext, void* xo, void* yo,
                                   void* zo, void* ho, size_t capacity, size_t* numLocal) = 0;
    virtual void view(hipcstone_pipeline_view* v) const = 0;
    virtual size_t deviceBytes() const = 0;
};

// curve-dispatched wrappers around the per-type C ABI
template<class K, class T>
struct Api;

#define HC_API(KS, K, RS, T)                                                                                           \
    template<>                                                                                                         \
    struct Api<K, T>                                                                                                   \
    {                                                                                                                  \
        using Box = hipcstone_box_##RS;                                                                                \
        static hipcstone_status keys(int curve, hipcstone_stream s, const T* x, const T* y, const T* z, K* k,         \
                                     size_t n, const Box* b)                                                           \
        {                                                                                                              \
            return curve == HIPCSTONE_HILBERT ? hipcstone_compute_sfc_keys_hilbert_##KS##_##RS(s, x, y, z, k, n, b, 0)\
                                              : hipcstone_compute_sfc_keys_morton_##KS##_##RS(s, x, y, z, k, n, b, 0);\
        }                                                                                                              \
        static size_t sortWs(size_t n) { return hipcstone_sort_gather_workspace_size_##KS(n); }                        \
        static hipcstone_status sortGather(hipcstone_stream s, K* k, uint32_t* p, size_t n, const T* xi, const T* yi, \
                                           const T* zi, const T* hi, T* xo, T* yo, T* zo, T* ho, K* kb, uint32_t* vb, \
                                           void* ws, size_t wsb)                                                       \
        {                                                                                                              \
            return hipcstone_sort_gather_##KS##_##RS(s, k, p, n, xi, yi, zi, hi, xo, yo, zo, ho, kb, vb, ws, wsb);    \
        }                                                                                                              \
        static size_t sortByKeyWs(size_t n) { return hipcstone_sort_by_key_workspace_size_##KS(n); }                  \
        static hipcstone_status sortByKey(hipcstone_stream s, K* k, uint32_t* v, size_t n, K* kb, uint32_t* vb,       \
                                          void* ws, size_t wsb)                                                        \
        {                                                                                                              \
            return hipcstone_sort_by_key_##KS(s, k, v, n, kb, vb, ws, wsb, int(8 * sizeof(K)));                       \
        }                                                                                                              \
        static size_t buildWs(int32_t L) { return hipcstone_build_octree_workspace_size_##KS(L); }                    \
        static hipcstone_status update(hipcstone_stream s, const K* keys, size_t n, uint32_t bucket, const K* tree,   \
                                       const uint32_t* counts, int32_t* nn, K* nt, uint32_t* nc, int32_t cap,          \
                                       void* ws, size_t wsb, int* conv)                                                \
        {                                                                                                              \
            return hipcstone_update_octree_##KS(s, keys, n, bucket, tree, counts, nn, nt, nc, cap, 0xFFFFFFFFu, ws,   \
                                                wsb, conv);                                                            \
        }                                                                                                              \
        static hipcstone_status build(hipcstone_stream s, const K* leaves, int32_t L, K* pf, int32_t* co,             \
                                      int32_t* par, int32_t* lr, int32_t* i2l, int32_t* l2i, void* ws, size_t wsb,     \
                                      int32_t* lrh)                                                                    \
        {                                                                                                              \
            return hipcstone_build_octree_##KS(s, leaves, L, pf, co, par, lr, i2l, l2i, ws, wsb, lrh);                \
        }                                                                                                              \
        static hipcstone_status centers(hipcstone_stream s, const K* pf, int32_t N, T* c, T* sz, const Box* b)        \
        {                                                                                                              \
            return hipcstone_compute_geo_centers_##KS##_##RS(s, pf, N, c, sz, b);                                     \
        }                                                                                                              \
        static hipcstone_status assign(hipcstone_stream s, const K* leaves, const uint32_t* counts, int32_t L,        \
                                       int32_t R, int32_t* off, K* bnd, uint32_t* rc, void* ws, size_t wsb)            \
        {                                                                                                              \
            return hipcstone_make_sfc_assignment_##KS(s, leaves, counts, L, R, off, bnd, rc, ws, wsb);                 \
        }                                                                                                              \
        static hipcstone_status discover(hipcstone_stream s, const K* pf, const int32_t* co, const int32_t* par,      \
                                         const int32_t* l2i, const T* c, const T* sz, const K* leaves, int32_t N,      \
                                         const T* x, const T* y, const T* z, const T* h, const uint32_t* lay,          \
                                         int32_t first, int32_t last, const Box* b, float ext, T* sc, T* ss,           \
                                         uint8_t* flags, void* ws, size_t wsb)                                         \
        {                                                                                                              \
            return hipcstone_discover_halos_##KS##_##RS(s, pf, co, par, l2i, c, sz, leaves, N, x, y, z, h, lay, first, \
                                                        last, b, ext, sc, ss, flags, 0, ws, wsb);                      \
        }                                                                                                              \
        static hipcstone_status gatherLeaves(hipcstone_stream s, const uint32_t* src, const uint32_t* dst, int32_t L, \
                                             const T* x, const T* y, const T* z, const T* h, T* xo, T* yo, T* zo,      \
                                             T* ho)                                                                    \
        {                                                                                                              \
            return hipcstone_gather_leaf_ranges_##RS(s, src, dst, L, x, y, z, h, xo, yo, zo, ho);                     \
        }                                                                                                              \
        static hipcstone_status neighbors(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h,         \
                                          uint32_t f, uint32_t l, const Box* b, const int32_t* co,                    \
                                          const int32_t* par, const int32_t* i2l, const uint32_t* lay, const T* c,     \
                                          const T* sz, int32_t N, uint32_t ngmax, uint32_t* nb, uint32_t* nc,          \
                                          void* gws, size_t gwsb)                                                      \
        {                                                                                                              \
            return findNeighborsWs<T>(s, x, y, z, h, f, l, b, co, par, i2l, lay, c, sz, N, 1.0f, ngmax, nb, nc, gws,  \
                                      gwsb);                                                                           \
        }                                                                                                              \
    };

HC_API(u32, uint32_t, f32, float)
HC_API(u32, uint32_t, f64, double)
HC_API(u64, uint64_t, f32, float)
HC_API(u64, uint64_t, f64, double)

// ---- direct construction of the converged leaf array (computeOctree, tree/csarray.hpp:429-440).
// The reference iterates updateOctree from the root until the input tree is converged. Its fixpoint is unique: a node
// is internal exactly when it holds more than bucketSize keys and lies above the deepest level (a node at or below
// the bucket size cannot be internal: its deepest internal descendant would have 8 leaf children whose total fits the
// bucket, which the rebalance merges; a node above it cannot be a leaf: the rebalance splits it). So the tree is built
// top down, one level per launch: every internal node of a level splits into its 8 children, whose key ranges come
// from binary searches inside the parent's range of the sorted keys; children above the bucket size (and above the
// deepest level) form the next level, the others are leaves. The leaves are then sorted by key, and the counts are the
// keys in each leaf's range, as computeNodeCounts gives them.
template<class K>
struct SplitNode
{
    K start;     // first key of the node
    uint32_t lo; // its range of the sorted keys
    uint32_t hi;
};

template<class K>
__device__ inline uint32_t lowerBoundIn(const K* __restrict__ keys, uint32_t lo, uint32_t hi, K v)
{
    while (lo < hi)
    {
        uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// the root's key range: keys at or past nodeRange(0) (removed particles) belong to no leaf
template<class K>
__global__ void splitRootKernel(const K* __restrict__ keys, uint32_t n, SplitNode<K>* __restrict__ frontier,
                                uint32_t* __restrict__ rootCount)
{
    uint32_t hi = lowerBoundIn(keys, 0u, n, nodeRange<K>(0));
    frontier[0] = SplitNode<K>{K(0), 0u, hi};
    *rootCount  = hi;
}

// 8 lanes per frontier node, lane k for child k; counters[0] counts the next frontier, counters[1] all leaves
template<class K>
__global__ __launch_bounds__(256) void splitLevelKernel(const K* __restrict__ keys, const SplitNode<K>* __restrict__ in,
                                                        uint32_t nIn, unsigned childLevel, uint32_t bucket,
                                                        SplitNode<K>* __restrict__ out, K* __restrict__ leafKeys,
                                                        uint32_t* __restrict__ leafCounts, uint32_t* counters)
{
    const uint32_t gid  = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t node = gid >> 3;
    const unsigned k    = gid & 7u;
    const bool valid    = node < nIn;
    SplitNode<K> f{K(0), 0u, 0u};
    if (valid) f = in[node];
    const K start = f.start + K(k) * nodeRange<K>(childLevel);
    uint32_t b    = (valid && k > 0) ? lowerBoundIn(keys, f.lo, f.hi, start) : f.lo;
    uint32_t e    = __shfl_down(b, 1, 8);
    if (k == 7) e = f.hi;
    const uint32_t cnt  = e - b;
    const bool internal = valid && cnt > bucket && childLevel < maxLevel<K>();
    const bool leaf     = valid && !internal;
    const uint64_t mi   = ballot(internal);
    const uint64_t ml   = ballot(leaf);
    uint32_t bi = 0, bl = 0;
    if (laneId() == 0)
    {
        if (mi) bi = atomicAdd(counters, uint32_t(__builtin_popcountll(mi)));
        if (ml) bl = atomicAdd(counters + 1, uint32_t(__builtin_popcountll(ml)));
    }
    bi = uniform(bi);
    bl = uniform(bl);
    if (internal) out[bi + laneRank(mi)] = SplitNode<K>{start, b, e};
    if (leaf)
    {
        leafKeys[bl + laneRank(ml)]   = start;
        leafCounts[bl + laneRank(ml)] = cnt;
    }
}

__global__ void setU32PairKernel(uint32_t* p, uint32_t a, uint32_t b)
{
    p[0] = a;
    p[1] = b;
}

template<class K, class T>
class Pipeline final : public PipelineBase
{
    using A   = Api<K, T>;
    using Box = typename A::Box;

public:
    Pipeline(int curve, uint32_t bucket, uint32_t ngmax)
        : curve_(curve)
        , bucket_(bucket)
        , ngmax_(ngmax)
    {
    }

    ~Pipeline() override
    {
        if (levelRangeReady_)
        {
            (void)hipEventSynchronize(levelRangeReady_);
            (void)hipEventDestroy(levelRangeReady_);
        }
        if (levelRangePinned_) (void)hipHostFree(levelRangePinned_);
    }

    hipcstone_status sync(hipcstone_stream s, const void* xv, const void* yv, const void* zv, const void* hv,
                          size_t n, const void* boxv, int fullBuild, uint32_t* nb, uint32_t* nc) override
    {
        if (n == 0) return invalid("pipeline sync: no particles");
        if (n > (size_t(1) << 30) - 1) return invalid("pipeline sync: more than 2^30-1 particles");
        if (!xv || !yv || !zv || !hv || !boxv) return invalid("pipeline sync: null input");
        const T* x = static_cast<const T*>(xv);
        const T* y = static_cast<const T*>(yv);
        const T* z = static_cast<const T*>(zv);
        const T* h = static_cast<const T*>(hv);
        box_       = *static_cast<const Box*>(boxv);

        if (auto e = reserveParticles(n)) return e;
        // gather into the particle set that does not alias the input
        int out = 1 - cur_;
        if (xyzh_[out][0].ptr == xv) out = cur_;
        T* xo = xyzh_[out][0].template as<T>();
        T* yo = xyzh_[out][1].template as<T>();
        T* zo = xyzh_[out][2].template as<T>();
        T* ho = xyzh_[out][3].template as<T>();

        K* keys = keys_.as<K>();
        auto& tm = this->timer;
#define HC_STAGE(id, call)                                                                                             \
    do {                                                                                                               \
        if (auto e_ = tm.begin(id, s)) return e_;                                                                     \
        if (auto e_ = (call)) return e_;                                                                               \
        if (auto e_ = tm.end(s)) return e_;                                                                            \
    } while (0)
        HC_STAGE(0, A::keys(curve_, s, x, y, z, keys, n, &box_));
        if (!fullBuild && numLeaves_ > 0) { HC_STAGE(1, sortByPreviousLeaves(s, x, y, z, h, xo, yo, zo, ho, n)); }
        else
        {
            HC_STAGE(1, A::sortGather(s, keys, perm_.as<uint32_t>(), n, x, y, z, h, xo, yo, zo, ho, keyBuf_.as<K>(),
                                      valBuf_.as<uint32_t>(), sortWs_.ptr, sortWs_.bytes));
        }
        cur_ = out;
        n_   = n;

        // the device's fault word (guards of the previous step's search and of this step's sort) rides along with
        // the synchronization of the tree build or update
        uint32_t* pin = pinnedStaging();
        if (!pin) return HIPCSTONE_HIP_ERROR;
        uint32_t* faults = pin + 8;
        *faults          = 0;
        if (auto e = enqueueFaultRead(stream(s), faults)) return e;
        if (fullBuild || numLeaves_ == 0)
        {
            HC_STAGE(2, computeOctree(s));
            lastMaxBucket_ = ~0u; // a new tree: the next warm step reads its largest bucket first
        }
        else
        {
            int conv = 0;
            HC_STAGE(2, updateOnce(s, &conv));
        }
        if (auto e = checkFaults(*faults)) return e;
        if (maxPending_)
        {
            lastMaxBucket_ = pin[9]; // the largest bucket of this step's sort, copied before the update synchronized
            maxPending_    = false;
        }
        HC_STAGE(3, link(s));
        HC_STAGE(4, A::centers(s, prefixes_.as<K>(), numNodes_, centers_.as<T>(), sizes_.as<T>(), &box_));
        HC_STAGE(5, hipcstone_leaf_layout(s, counts_[treeCur_].as<uint32_t>(), numLeaves_, layout_.as<uint32_t>(),
                                          scanWs_.ptr, scanWs_.bytes));
        if (nc)
        {
            if (auto e = groupsWs_.reserve(neighborGroupsBytes(numNodes_))) return e;
            HC_STAGE(6, A::neighbors(s, xo, yo, zo, ho, 0, uint32_t(n), &box_, childOffsets_.as<int32_t>(),
                                     parents_.as<int32_t>(), internalToLeaf_.as<int32_t>(), layout_.as<uint32_t>(),
                                     centers_.as<T>(), sizes_.as<T>(), numNodes_, nb ? ngmax_ : 0u, nb, nc,
                                     groupsWs_.ptr, groupsWs_.bytes));
        }
#undef HC_STAGE
        return HIPCSTONE_OK;
    }

    hipcstone_status halos(hipcstone_stream s, int32_t numRanks, int32_t rank, float ext, void* xo, void* yo,
                           void* zo, void* ho, size_t capacity, size_t* numLocal) override
    {
        if (numLeaves_ == 0 || numNodes_ == 0) return invalid("pipeline halos: no synced tree");
        if (numRanks <= 0 || rank < 0 || rank >= numRanks) return invalid("pipeline halos: bad rank");
        if (!numLocal) return invalid("pipeline halos: null num_local");
        const int32_t L = numLeaves_, N = numNodes_;
        size_t R = size_t(numRanks);
        if (auto e = assignOff_.reserve((R + 1) * sizeof(int32_t))) return e;
        if (auto e = assignBnd_.reserve((R + 1) * sizeof(K))) return e;
        if (auto e = assignCnt_.reserve(R * sizeof(uint32_t))) return e;
        if (auto e = assignWs_.reserve(hipcstone_sfc_assignment_workspace_size(L))) return e;
        if (auto e = flags_.reserve(size_t(N))) return e;
        if (auto e = searchC_.reserve(size_t(L) * 3 * sizeof(T))) return e;
        if (auto e = searchS_.reserve(size_t(L) * 3 * sizeof(T))) return e;
        if (auto e = localLayout_.reserve((size_t(L) + 1) * sizeof(uint32_t))) return e;
        const int32_t* leafToInternalLeaves = leafToInternal_.as<int32_t>() + (N - L); // tree/octree.hpp:252
        const T* x = xyzh_[cur_][0].template as<T>();
        const T* y = xyzh_[cur_][1].template as<T>();
        const T* z = xyzh_[cur_][2].template as<T>();
        const T* h = xyzh_[cur_][3].template as<T>();
        auto& tm = this->timer;
#define HC_STAGE(id, call)                                                                                             \
    do {                                                                                                               \
        if (auto e_ = tm.begin(id, s)) return e_;                                                                     \
        if (auto e_ = (call)) return e_;                                                                               \
        if (auto e_ = tm.end(s)) return e_;                                                                            \
    } while (0)
        HC_STAGE(7, A::assign(s, tree_[treeCur_].template as<K>(), counts_[treeCur_].template as<uint32_t>(), L,
                              numRanks, assignOff_.as<int32_t>(), assignBnd_.as<K>(), assignCnt_.as<uint32_t>(),
                              assignWs_.ptr, assignWs_.bytes));
        int32_t range[2];
        HC_CHECK(hipMemcpyAsync(range, assignOff_.as<int32_t>() + rank, sizeof(range), hipMemcpyDeviceToHost,
                                stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s)));
        first_ = range[0];
        last_  = range[1];
        if (auto e = haloWs_.reserve(hipcstone_find_halos_workspace_size(N, last_ - first_, int32_t(sizeof(T)))))
            return e;
        HC_STAGE(8, A::discover(s, prefixes_.as<K>(), childOffsets_.as<int32_t>(), parents_.as<int32_t>(),
                                leafToInternalLeaves, centers_.as<T>(), sizes_.as<T>(), tree_[treeCur_].template as<K>(),
                                N, x, y, z, h, layout_.as<uint32_t>(), first_, last_, &box_, ext, searchC_.as<T>(),
                                searchS_.as<T>(), flags_.as<uint8_t>(), haloWs_.ptr, haloWs_.bytes));
        HC_STAGE(9, hipcstone_compute_node_layout(s, counts_[treeCur_].template as<uint32_t>(), flags_.as<uint8_t>(),
                                                  leafToInternalLeaves, L, first_, last_, localLayout_.as<uint32_t>(),
                                                  scanWs_.ptr, scanWs_.bytes));
        uint32_t total = 0;
        HC_CHECK(hipMemcpyAsync(&total, localLayout_.as<uint32_t>() + L, sizeof(total), hipMemcpyDeviceToHost,
                                stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s)));
        numLocal_ = total;
        *numLocal = total;
        if (!xo && !yo && !zo && !ho) return HIPCSTONE_OK;
        if (!xo || !yo || !zo || !ho) return invalid("pipeline halos: give all four outputs or none");
        if (capacity < total)
        {
            setError("pipeline halos: output capacity below the local particle count");
            return HIPCSTONE_CAPACITY;
        }
        HC_STAGE(10, A::gatherLeaves(s, layout_.as<uint32_t>(), localLayout_.as<uint32_t>(), L, x, y, z, h,
                                     static_cast<T*>(xo), static_cast<T*>(yo), static_cast<T*>(zo),
                                     static_cast<T*>(ho)));
#undef HC_STAGE
        return HIPCSTONE_OK;
    }

    void view(hipcstone_pipeline_view* v) const override
    {
        v->num_particles    = n_;
        v->num_leaves       = numLeaves_;
        v->num_nodes        = numNodes_;
        v->num_internal     = numNodes_ - numLeaves_;
        v->keys             = keys_.ptr;
        v->perm             = perm_.ptr;
        v->x                = xyzh_[cur_][0].ptr;
        v->y                = xyzh_[cur_][1].ptr;
        v->z                = xyzh_[cur_][2].ptr;
        v->h                = xyzh_[cur_][3].ptr;
        v->leaves           = tree_[treeCur_].ptr;
        v->counts           = counts_[treeCur_].ptr;
        v->prefixes         = prefixes_.ptr;
        v->child_offsets    = childOffsets_.ptr;
        v->parents          = parents_.ptr;
        v->internal_to_leaf = internalToLeaf_.ptr;
        v->leaf_to_internal = leafToInternal_.ptr;
        v->level_range      = levelRange_.ptr;
        v->centers          = centers_.ptr;
        v->sizes            = sizes_.ptr;
        v->layout           = layout_.ptr;
        // the level range's copy of the last link (an asynchronous copy into pinned memory: wait for it)
        if (levelRangeReady_) (void)hipEventSynchronize(levelRangeReady_);
        for (int i = 0; i < 24; ++i)
            v->level_range_host[i] = levelRangePinned_ && i < int(maxLevel<K>() + 2) ? levelRangePinned_[i] : 0;
        v->converged_calls = convergedCalls_;
        v->assigned_first  = first_;
        v->assigned_last   = last_;
        v->num_local       = numLocal_;
        v->tree_offsets    = assignOff_.ptr;
        v->halo_flags      = flags_.ptr;
        v->search_centers  = searchC_.ptr;
        v->search_sizes    = searchS_.ptr;
        v->local_layout    = localLayout_.ptr;
    }

    // device memory held by the pipeline's buffers
    size_t deviceBytes() const override
    {
        size_t b = 0;
        for (const DeviceBuf* d : {&keys_, &keyBuf_, &perm_, &valBuf_, &sortWs_, &segWs_, &hintWs_, &updWs_, &posWs_,
                                   &splitKeys_, &splitCounts_, &splitCounters_, &splitSortWs_, &prefixes_, &childOffsets_,
                                   &parents_, &levelRange_, &internalToLeaf_, &leafToInternal_, &buildWs_, &centers_,
                                   &sizes_, &layout_, &scanWs_, &assignOff_, &assignBnd_, &assignCnt_, &assignWs_,
                                   &flags_, &searchC_, &searchS_, &localLayout_, &haloWs_, &groupsWs_})
            b += d->bytes;
        for (int t = 0; t < 2; ++t)
        {
            b += tree_[t].bytes + counts_[t].bytes + frontier_[t].bytes;
            for (int a = 0; a < 4; ++a)
                b += xyzh_[t][a].bytes;
        }
        return b;
    }

private:
    // warm-step stable sort + gather through the buckets of the previous leaf array (numLeaves_ > 0); the radix
    // sort when a bucket is too large
    hipcstone_status sortByPreviousLeaves(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, T* xo,
                                          T* yo, T* zo, T* ho, size_t n)
    {
        bool done = false;
        // the previous step's layout is over the previous positions of the same particles when their count is
        // unchanged (a search hint only: any order of the input gives the same result)
        const uint32_t* prevLayout = n == n_ ? layout_.as<uint32_t>() : nullptr;
        // when the previous step's largest bucket was well below the guard, this step's is not read before the
        // bucket path runs: it comes back with the tree update's synchronization (pinned word 9) and guards the
        // next step
        uint32_t* pin = pinnedStaging();
        if (!pin) return HIPCSTONE_HIP_ERROR;
        const bool defer = lastMaxBucket_ <= kMaxBucket / 4;
        uint32_t largest = ~0u;
        uint32_t* maxOut = defer ? pin + 9 : &largest;
        *maxOut          = ~0u;
        if (auto e = sortByLeafBuckets<K, T>(s, keys_.as<K>(), n, tree_[treeCur_].template as<K>(), numLeaves_, x, y,
                                             z, h, xo, yo, zo, ho, perm_.as<uint32_t>(), keyBuf_.as<K>(),
                                             valBuf_.as<uint32_t>(), segWs_, scanWs_, &done, prevLayout, &hintWs_,
                                             maxOut, defer))
            return e;
        maxPending_    = defer;
        lastMaxBucket_ = defer ? ~0u : largest;
        warmPos_       = done; // the bucket offsets are the old boundaries' positions for the tree update
        if (done) return HIPCSTONE_OK;
        return A::sortGather(s, keys_.as<K>(), perm_.as<uint32_t>(), n, x, y, z, h, xo, yo, zo, ho, keyBuf_.as<K>(),
                             valBuf_.as<uint32_t>(), sortWs_.ptr, sortWs_.bytes);
    }

    hipcstone_status reserveParticles(size_t n)
    {
        if (n <= capN_) return HIPCSTONE_OK;
        for (auto* b : {&keys_, &keyBuf_})
            if (auto e = b->reserve(n * sizeof(K))) return e;
        for (auto* b : {&perm_, &valBuf_})
            if (auto e = b->reserve(n * sizeof(uint32_t))) return e;
        for (int s = 0; s < 2; ++s)
            for (int a = 0; a < 4; ++a)
                if (auto e = xyzh_[s][a].reserve(n * sizeof(T))) return e;
        if (auto e = sortWs_.reserve(A::sortWs(n))) return e;
        capN_ = n;
        return HIPCSTONE_OK;
    }

    // leaf-array buffers for `cap` leaves (tree ping-pong, counts, update workspace)
    hipcstone_status reserveLeaves(int32_t cap)
    {
        if (cap <= capLeaves_) return HIPCSTONE_OK;
        // keep the current tree across the reallocation
        std::vector<K> keepTree;
        std::vector<uint32_t> keepCounts;
        if (numLeaves_ > 0)
        {
            keepTree.resize(numLeaves_ + 1);
            keepCounts.resize(numLeaves_);
            HC_CHECK(hipMemcpy(keepTree.data(), tree_[treeCur_].ptr, keepTree.size() * sizeof(K),
                               hipMemcpyDeviceToHost));
            HC_CHECK(hipMemcpy(keepCounts.data(), counts_[treeCur_].ptr, keepCounts.size() * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
        }
        int32_t c = std::max(cap, capLeaves_ * 2);
        for (int t = 0; t < 2; ++t)
        {
            if (auto e = tree_[t].reserve((size_t(c) + 1) * sizeof(K))) return e;
            if (auto e = counts_[t].reserve(size_t(c) * sizeof(uint32_t))) return e;
        }
        if (auto e = updWs_.reserve(hipcstone_update_octree_workspace_size(c))) return e;
        if (numLeaves_ > 0)
        {
            HC_CHECK(hipMemcpy(tree_[treeCur_].ptr, keepTree.data(), keepTree.size() * sizeof(K),
                               hipMemcpyHostToDevice));
            HC_CHECK(hipMemcpy(counts_[treeCur_].ptr, keepCounts.data(), keepCounts.size() * sizeof(uint32_t),
                               hipMemcpyHostToDevice));
        }
        capLeaves_ = c;
        return HIPCSTONE_OK;
    }

    // one updateOctree step from (tree_[treeCur_], counts_[treeCur_]); grows buffers and retries on CAPACITY
    hipcstone_status updateOnce(hipcstone_stream s, int* converged)
    {
        // after a bucket sort through the same leaves, the node counts come from its bucket offsets
        const bool warm = warmPos_;
        warmPos_        = false;
        for (int attempt = 0; attempt < 8; ++attempt)
        {
            int32_t nn = numLeaves_;
            int other  = 1 - treeCur_;
            hipcstone_status e;
            if (warm)
            {
                if (auto g = posWs_.reserve((size_t(capLeaves_) + 1) * sizeof(uint32_t))) return g;
                e = updateOctreeWarm<K>(s, keys_.as<K>(), n_, bucket_, tree_[treeCur_].template as<K>(),
                                        counts_[treeCur_].template as<uint32_t>(), &nn, tree_[other].template as<K>(),
                                        counts_[other].template as<uint32_t>(), capLeaves_, updWs_.ptr, updWs_.bytes,
                                        converged, segWs_.as<uint32_t>() + numLeaves_, posWs_.as<uint32_t>());
            }
            else
            {
                e = A::update(s, keys_.as<K>(), n_, bucket_, tree_[treeCur_].template as<K>(),
                              counts_[treeCur_].template as<uint32_t>(), &nn, tree_[other].template as<K>(),
                              counts_[other].template as<uint32_t>(), capLeaves_, updWs_.ptr, updWs_.bytes,
                              converged);
            }
            if (e == HIPCSTONE_CAPACITY)
            {
                if (auto g = reserveLeaves(nn)) return g;
                continue;
            }
            if (e) return e;
            treeCur_   = other;
            numLeaves_ = nn;
            return HIPCSTONE_OK;
        }
        return invalid("pipeline: leaf capacity growth did not converge");
    }

    // computeOctree: the converged leaf array of the sorted keys, built level by level (splitLevelKernel)
    hipcstone_status computeOctree(hipcstone_stream s)
    {
        numLeaves_ = 0; // nothing to keep across buffer growth
        warmPos_   = false;
        if (auto e = reserveLeaves(std::max<int32_t>(capLeaves_, 4096))) return e;
        if (auto e = splitCounters_.reserve(2 * sizeof(uint32_t))) return e;
        if (auto e = frontier_[0].reserve(sizeof(SplitNode<K>))) return e;
        splitRootKernel<K><<<1, 1, 0, stream(s)>>>(keys_.as<K>(), uint32_t(n_), frontier_[0].as<SplitNode<K>>(),
                                                   splitCounters_.as<uint32_t>());
        HC_CHECK_LAUNCH("splitRootKernel");
        uint32_t host[2];
        HC_CHECK(hipMemcpyAsync(host, splitCounters_.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s)));
        const K rootKeys[2] = {K(0), nodeRange<K>(0)};
        if (host[0] <= bucket_)
        {
            HC_CHECK(hipMemcpyAsync(tree_[treeCur_].ptr, rootKeys, sizeof(rootKeys), hipMemcpyHostToDevice,
                                    stream(s)));
            HC_CHECK(hipMemcpyAsync(counts_[treeCur_].ptr, &host[0], sizeof(uint32_t), hipMemcpyHostToDevice,
                                    stream(s)));
            numLeaves_      = 1;
            convergedCalls_ = 0;
            return HIPCSTONE_OK;
        }
        uint32_t nIn = 1, leaves = 0;
        int cur = 0, levels = 0;
        for (unsigned childLevel = 1; nIn > 0; ++childLevel)
        {
            if (childLevel > maxLevel<K>()) return invalid("pipeline: computeOctree split below the deepest level");
            const size_t nOutMax = size_t(nIn) * 8;
            if (auto e = frontier_[1 - cur].reserve(nOutMax * sizeof(SplitNode<K>))) return e;
            if (auto e = growKeep(splitKeys_, (leaves + nOutMax) * sizeof(K), leaves * sizeof(K), s)) return e;
            if (auto e = growKeep(splitCounts_, (leaves + nOutMax) * sizeof(uint32_t), leaves * sizeof(uint32_t), s))
                return e;
            setU32PairKernel<<<1, 1, 0, stream(s)>>>(splitCounters_.as<uint32_t>(), 0u, leaves);
            HC_CHECK_LAUNCH("setU32PairKernel");
            splitLevelKernel<K><<<ceilDiv(nOutMax, 256), 256, 0, stream(s)>>>(
                keys_.as<K>(), frontier_[cur].as<SplitNode<K>>(), nIn, childLevel, bucket_,
                frontier_[1 - cur].as<SplitNode<K>>(), splitKeys_.as<K>(), splitCounts_.as<uint32_t>(),
                splitCounters_.as<uint32_t>());
            HC_CHECK_LAUNCH("splitLevelKernel");
            HC_CHECK(hipMemcpyAsync(host, splitCounters_.ptr, sizeof(host), hipMemcpyDeviceToHost, stream(s)));
            HC_CHECK(hipStreamSynchronize(stream(s)));
            nIn    = host[0];
            leaves = host[1];
            cur    = 1 - cur;
            ++levels;
        }
        if (leaves > uint32_t(INT32_MAX - 1)) return invalid("pipeline: computeOctree leaf count above 2^31");
        const int32_t L = int32_t(leaves);
        if (auto e = reserveLeaves(L)) return e;
        K* tree          = tree_[treeCur_].template as<K>();
        uint32_t* counts = counts_[treeCur_].template as<uint32_t>();
        HC_CHECK(hipMemcpyAsync(tree, splitKeys_.ptr, size_t(L) * sizeof(K), hipMemcpyDeviceToDevice, stream(s)));
        HC_CHECK(hipMemcpyAsync(counts, splitCounts_.ptr, size_t(L) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                                stream(s)));
        // leaf keys are distinct, so any sort orders them the same way; the buffers of the splits serve as scratch
        if (auto e = splitSortWs_.reserve(A::sortByKeyWs(size_t(L)))) return e;
        if (auto e = A::sortByKey(s, tree, counts, size_t(L), splitKeys_.as<K>(), splitCounts_.as<uint32_t>(),
                                  splitSortWs_.ptr, splitSortWs_.bytes))
            return e;
        HC_CHECK(hipMemcpyAsync(tree + L, &rootKeys[1], sizeof(K), hipMemcpyHostToDevice, stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s))); // rootKeys lives on this stack frame
        numLeaves_      = L;
        convergedCalls_ = levels;
        return HIPCSTONE_OK;
    }

    // grow a device buffer to `bytes`, keeping its first `keep` bytes
    static hipcstone_status growKeep(DeviceBuf& b, size_t bytes, size_t keep, hipcstone_stream s)
    {
        if (bytes <= b.bytes) return HIPCSTONE_OK;
        DeviceBuf nb;
        if (auto e = nb.reserve(std::max(bytes, 2 * b.bytes))) return e;
        if (keep) HC_CHECK(hipMemcpyAsync(nb.ptr, b.ptr, keep, hipMemcpyDeviceToDevice, stream(s)));
        HC_CHECK(hipStreamSynchronize(stream(s))); // the old buffer is freed below
        swap(b, nb);
        return HIPCSTONE_OK;
    }

    hipcstone_status link(hipcstone_stream s)
    {
        int32_t L = numLeaves_;
        int32_t N = L + (L - 1) / 7;
        if (N > capNodes_)
        {
            int32_t c = std::max(N, capNodes_ * 2);
            size_t Nc = size_t(c);
            if (auto e = prefixes_.reserve(Nc * sizeof(K))) return e;
            for (auto* b : {&childOffsets_, &internalToLeaf_, &leafToInternal_})
                if (auto e = b->reserve(Nc * sizeof(int32_t))) return e;
            if (auto e = parents_.reserve(std::max<size_t>(1, (Nc - 1) / 8 + 1) * sizeof(int32_t))) return e;
            if (auto e = levelRange_.reserve(32 * sizeof(int32_t))) return e;
            if (auto e = centers_.reserve(Nc * 3 * sizeof(T))) return e;
            if (auto e = sizes_.reserve(Nc * 3 * sizeof(T))) return e;
            capNodes_ = c;
        }
        if (auto e = buildWs_.reserve(A::buildWs(L))) return e;
        if (auto e = layout_.reserve((size_t(L) + 1) * sizeof(uint32_t))) return e;
        if (auto e = scanWs_.reserve(hipcstone_scan_workspace_size(size_t(L)))) return e;
        numNodes_ = N;
        if (auto e = A::build(s, tree_[treeCur_].template as<K>(), L, prefixes_.as<K>(), childOffsets_.as<int32_t>(),
                              parents_.as<int32_t>(), levelRange_.as<int32_t>(), internalToLeaf_.as<int32_t>(),
                              leafToInternal_.as<int32_t>(), buildWs_.ptr, buildWs_.bytes, nullptr))
            return e;
        // the view's host copy of the level range: an asynchronous copy into pinned memory, waited for by view()
        if (!levelRangePinned_)
        {
            void* q = nullptr;
            HC_CHECK(hipHostMalloc(&q, 32 * sizeof(int32_t), hipHostMallocDefault));
            levelRangePinned_ = static_cast<int32_t*>(q);
            HC_CHECK(hipEventCreateWithFlags(&levelRangeReady_, hipEventDisableTiming));
        }
        HC_CHECK(hipMemcpyAsync(levelRangePinned_, levelRange_.ptr, sizeof(int32_t) * (maxLevel<K>() + 2),
                                hipMemcpyDeviceToHost, stream(s)));
        HC_CHECK(hipEventRecord(levelRangeReady_, stream(s)));
        return HIPCSTONE_OK;
    }

    int curve_;
    uint32_t bucket_;
    uint32_t ngmax_;
    Box box_{};
    size_t n_          = 0;
    size_t capN_       = 0;
    int32_t capLeaves_ = 0;
    int32_t capNodes_  = 0;
    int32_t numLeaves_ = 0;
    int32_t numNodes_  = 0;
    int cur_           = 0;
    int treeCur_       = 0;
    int convergedCalls_ = 0;
    uint32_t lastMaxBucket_     = ~0u;   // largest bucket of the last warm-step sort, ~0 when unknown
    bool warmPos_               = false; // this step's bucket sort left the old boundaries' positions in segWs_
    bool maxPending_            = false; // this step's largest bucket is on its way to pinned word 9
    int32_t* levelRangePinned_  = nullptr; // hipHostMalloc'd
    hipEvent_t levelRangeReady_ = nullptr;
    int32_t first_     = 0;
    int32_t last_      = 0;
    size_t numLocal_   = 0;

    DeviceBuf keys_, keyBuf_, perm_, valBuf_, sortWs_, segWs_, hintWs_;
    DeviceBuf xyzh_[2][4];
    DeviceBuf tree_[2], counts_[2], updWs_, posWs_;
    DeviceBuf frontier_[2], splitKeys_, splitCounts_, splitCounters_, splitSortWs_;
    DeviceBuf prefixes_, childOffsets_, parents_, levelRange_, internalToLeaf_, leafToInternal_, buildWs_;
    DeviceBuf centers_, sizes_, layout_, scanWs_, groupsWs_;
    DeviceBuf assignOff_, assignBnd_, assignCnt_, assignWs_, flags_, searchC_, searchS_, localLayout_, haloWs_;
};

} // namespace hipcstone

struct hipcstone_pipeline
{
    std::unique_ptr<hipcstone::PipelineBase> impl;
};

using namespace hipcstone;

extern "C"
{

hipcstone_status hipcstone_pipeline_create(hipcstone_pipeline** p, int curve, int key_bits, int real_bits,
                                           uint32_t bucket_size, uint32_t ngmax)
{
    if (!p) return invalid("pipeline_create: null handle");
    if (curve != HIPCSTONE_HILBERT && curve != HIPCSTONE_MORTON) return invalid("pipeline_create: curve");
    if (bucket_size == 0) return invalid("pipeline_create: bucket_size must be positive");
    auto* h = new hipcstone_pipeline;
    if (key_bits == 64 && real_bits == 32) h->impl = std::make_unique<Pipeline<uint64_t, float>>(curve, bucket_size, ngmax);
    else if (key_bits == 64 && real_bits == 64) h->impl = std::make_unique<Pipeline<uint64_t, double>>(curve, bucket_size, ngmax);
    else if (key_bits == 32 && real_bits == 32) h->impl = std::make_unique<Pipeline<uint32_t, float>>(curve, bucket_size, ngmax);
    else if (key_bits == 32 && real_bits == 64) h->impl = std::make_unique<Pipeline<uint32_t, double>>(curve, bucket_size, ngmax);
    else
    {
        delete h;
        return invalid("pipeline_create: key_bits must be 32|64 and real_bits 32|64");
    }
    *p = h;
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_pipeline_destroy(hipcstone_pipeline* p)
{
    delete p;
    return HIPCSTONE_OK;
}

size_t hipcstone_pipeline_device_bytes(const hipcstone_pipeline* p) { return p ? p->impl->deviceBytes() : 0; }

hipcstone_status hipcstone_pipeline_sync(hipcstone_pipeline* p, hipcstone_stream s, const void* x, const void* y,
                                         const void* z, const void* h, size_t n, const void* box, int full_build,
                                         uint32_t* neighbors, uint32_t* neighbors_count)
{
    if (!p) return invalid("pipeline_sync: null handle");
    return p->impl->sync(s, x, y, z, h, n, box, full_build, neighbors, neighbors_count);
}

hipcstone_status hipcstone_pipeline_halos(hipcstone_pipeline* p, hipcstone_stream s, int32_t num_ranks, int32_t rank,
                                          float search_ext, void* x_out, void* y_out, void* z_out, void* h_out,
                                          size_t capacity, size_t* num_local)
{
    if (!p) return invalid("pipeline_halos: null handle");
    return p->impl->halos(s, num_ranks, rank, search_ext, x_out, y_out, z_out, h_out, capacity, num_local);
}

hipcstone_status hipcstone_pipeline_set_timing(hipcstone_pipeline* p, int enable)
{
    if (!p) return invalid("pipeline_set_timing: null handle");
    p->impl->timer.enable(enable != 0);
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_pipeline_stage_times(hipcstone_pipeline* p, double* ms, int32_t* launches)
{
    if (!p) return invalid("pipeline_stage_times: null handle");
    return p->impl->timer.collect(ms, launches);
}

hipcstone_status hipcstone_pipeline_get_view(const hipcstone_pipeline* p, hipcstone_pipeline_view* v)
{
    if (!p || !v) return invalid("pipeline_get_view: null pointer");
    p->impl->view(v);
    return HIPCSTONE_OK;
}

} // extern "C"
