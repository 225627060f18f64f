/*
 * ORACLE — test infrastructure, not product code.
 *
 * C-callable driver of the reference's own one-rank Domain (domain/domain.hpp:38-420): the reference headers are
 * included where they lie (-I /root/reference/include, see oracle/Makefile target `refdomain`) and built into
 * oracle/_ref/libcstone_ref_domain.so, linked against the image's MPI library and run as an MPI singleton (one
 * rank, no launcher). Only tests/golden/make_golden.py and tests/ load it, to pin the oracle's one-rank Domain
 * sequence (oracle_py.CpuDomain) and to write the domain golden fixtures. It contains no reference code.
 *
 * Every function takes the opaque handle returned by refdom_create. Arrays are host arrays; T and K are fixed by
 * the create call (real_bits 32/64, key_bits 32/64). The Domain always uses Hilbert keys (domain.hpp:430-447,
 * SfcKind of the KeyType template argument is HilbertKey through sfcKindPointer in assignment.hpp).
 */
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include <mpi.h>

#include "cstone/domain/domain.hpp"

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

#pragma GCC diagnostic ignored "-Wnon-template-friend"
// read access to Domain::global_ (the global leaf array and its counts have no public accessor on Domain): the
// address of a private member named in an explicit instantiation is exempt from access checking
template<class Tag, typename Tag::type M>
struct Expose
{
    friend typename Tag::type memberOf(Tag) { return M; }
};

template<class K, class T>
struct GlobalTag
{
    using D    = Domain<K, T, execution::Cpu>;
    using type = GlobalAssignment<K, T, execution::Cpu> D::*;
    friend type memberOf(GlobalTag);
};

template struct Expose<GlobalTag<uint32_t, float>, &Domain<uint32_t, float, execution::Cpu>::global_>;
template struct Expose<GlobalTag<uint32_t, double>, &Domain<uint32_t, double, execution::Cpu>::global_>;
template struct Expose<GlobalTag<uint64_t, float>, &Domain<uint64_t, float, execution::Cpu>::global_>;
template struct Expose<GlobalTag<uint64_t, double>, &Domain<uint64_t, double, execution::Cpu>::global_>;

std::string lastError;

void ensureMpi()
{
    int init = 0;
    MPI_Initialized(&init);
    if (!init) MPI_Init(nullptr, nullptr);
}

struct Base
{
    virtual ~Base() = default;
    virtual int sync(void* x, void* y, void* z, void* h, void* m, uint32_t* ids, size_t n, void* keysOut) = 0;
    virtual void sizes(int64_t* out) const                                                           = 0;
    virtual void box(double* out) const                                                              = 0;
    virtual void global(void* leaves, uint32_t* counts) const                                        = 0;
    virtual void focus(void* leaves, uint32_t* leafCounts, uint32_t* nodeCounts, uint32_t* layout) const = 0;
    virtual void octree(void* prefixes, int32_t* childOffsets, int32_t* internalToLeaf, int32_t* levelRange,
                        void* centers, void* sizes) const                                            = 0;
    virtual void expansion(void* centers) const                                                      = 0;
};

template<class K, class T>
struct Dom final : Base
{
    using D = Domain<K, T, execution::Cpu>;
    std::unique_ptr<D> d;
    std::vector<K> keys;
    std::vector<T> vx, vy, vz, vh, vm;
    std::vector<uint32_t> vid;
    bool grav = false;

    Dom(unsigned bucket, unsigned bucketFocus, float theta, const BoxC<T>* b)
    {
        Box<T> box(b->lim[0], b->lim[1], b->lim[2], b->lim[3], b->lim[4], b->lim[5], BoundaryType(b->bnd[0]),
                   BoundaryType(b->bnd[1]), BoundaryType(b->bnd[2]));
        d = std::make_unique<D>(execution::cpu, 0, 1, bucket, bucketFocus, theta, MPI_COMM_WORLD, box);
    }

    const GlobalAssignment<K, T, execution::Cpu>& globalAssignment() const
    {
        return (*d).*memberOf(GlobalTag<K, T>{});
    }

    // x, y, z, h (and m) are updated in place; ids, when given, receives the input index of every output
    // particle (carried through the sync as a particle property)
    int sync(void* x, void* y, void* z, void* h, void* m, uint32_t* ids, size_t n, void* keysOut) override
    {
        auto load = [n](std::vector<T>& v, const void* p) { v.assign((const T*)p, (const T*)p + n); };
        load(vx, x);
        load(vy, y);
        load(vz, z);
        load(vh, h);
        vid.resize(n);
        std::iota(vid.begin(), vid.end(), 0u);
        keys.resize(n);
        std::vector<T> s1, s2, s3;
        grav = m != nullptr;
        if (grav)
        {
            load(vm, m);
            d->syncGrav(keys, vx, vy, vz, vh, vm, std::tie(vid), std::tie(s1, s2, s3));
        }
        else { d->sync(keys, vx, vy, vz, vh, std::tie(vid), std::tie(s1, s2, s3)); }
        if (vx.size() != n) return -1; // one rank has no halos: the sizes must not change
        std::memcpy(x, vx.data(), n * sizeof(T));
        std::memcpy(y, vy.data(), n * sizeof(T));
        std::memcpy(z, vz.data(), n * sizeof(T));
        std::memcpy(h, vh.data(), n * sizeof(T));
        if (grav) std::memcpy(m, vm.data(), n * sizeof(T));
        if (ids) std::memcpy(ids, vid.data(), n * sizeof(uint32_t));
        if (keysOut) std::memcpy(keysOut, keys.data(), n * sizeof(K));
        return 0;
    }

    // [start, end, nParticlesWithHalos, global leaves, focus leaves, focus nodes, focus internal nodes]
    void sizes(int64_t* out) const override
    {
        auto ov = d->focusTree().octreeViewAcc();
        out[0]  = d->startIndex();
        out[1]  = d->endIndex();
        out[2]  = d->nParticlesWithHalos();
        out[3]  = int64_t(globalAssignment().treeLeaves().size()) - 1;
        out[4]  = ov.numLeafNodes;
        out[5]  = ov.numNodes;
        out[6]  = ov.numInternalNodes;
    }

    void box(double* out) const override
    {
        const Box<T>& b = d->box();
        out[0] = b.xmin(), out[1] = b.xmax(), out[2] = b.ymin(), out[3] = b.ymax(), out[4] = b.zmin();
        out[5] = b.zmax();
    }

    void global(void* leaves, uint32_t* counts) const override
    {
        auto l = globalAssignment().treeLeaves();
        auto c = globalAssignment().nodeCounts();
        std::memcpy(leaves, l.data(), l.size() * sizeof(K));
        std::memcpy(counts, c.data(), c.size() * sizeof(unsigned));
    }

    void focus(void* leaves, uint32_t* leafCounts, uint32_t* nodeCounts, uint32_t* layout) const override
    {
        auto& f = d->focusTree();
        auto l  = f.treeLeaves();
        auto lc = f.leafCountsAcc();
        auto nc = f.countsAcc();
        auto ly = d->layout();
        std::memcpy(leaves, l.data(), l.size() * sizeof(K));
        std::memcpy(leafCounts, lc.data(), lc.size() * sizeof(unsigned));
        std::memcpy(nodeCounts, nc.data(), nc.size() * sizeof(unsigned));
        std::memcpy(layout, ly.data(), ly.size() * sizeof(LocalIndex));
    }

    void octree(void* prefixes, int32_t* childOffsets, int32_t* internalToLeaf, int32_t* levelRange, void* centers,
                void* sizes) const override
    {
        auto ov = d->focusTree().octreeViewAcc();
        size_t N = ov.numNodes;
        std::memcpy(prefixes, ov.prefixes, N * sizeof(K));
        std::memcpy(childOffsets, ov.childOffsets, N * sizeof(int32_t));
        std::memcpy(internalToLeaf, ov.internalToLeaf, N * sizeof(int32_t));
        std::memcpy(levelRange, ov.levelRange, (maxTreeLevel<K>{} + 2) * sizeof(int32_t));
        auto gc = d->focusTree().geoCentersAcc();
        auto gs = d->focusTree().geoSizesAcc();
        std::memcpy(centers, gc.data(), N * 3 * sizeof(T));
        std::memcpy(sizes, gs.data(), N * 3 * sizeof(T));
    }

    void expansion(void* centers) const override
    {
        auto c = d->focusTree().expansionCentersAcc();
        std::memcpy(centers, c.data(), c.size() * 4 * sizeof(T));
    }
};

} // namespace

extern "C"
{

// box: a BoxC<float> or BoxC<double> matching real_bits (the layout of hipcstone_box_f32/f64)
void* refdom_create(int key_bits, int real_bits, unsigned bucket, unsigned bucket_focus, float theta, const void* box)
{
    ensureMpi();
    try
    {
        if (key_bits == 32 && real_bits == 32)
            return new Dom<uint32_t, float>(bucket, bucket_focus, theta, (const BoxC<float>*)box);
        if (key_bits == 32 && real_bits == 64)
            return new Dom<uint32_t, double>(bucket, bucket_focus, theta, (const BoxC<double>*)box);
        if (key_bits == 64 && real_bits == 32)
            return new Dom<uint64_t, float>(bucket, bucket_focus, theta, (const BoxC<float>*)box);
        if (key_bits == 64 && real_bits == 64)
            return new Dom<uint64_t, double>(bucket, bucket_focus, theta, (const BoxC<double>*)box);
    }
    catch (...)
    {
    }
    return nullptr;
}

const char* refdom_last_error(void) { return lastError.c_str(); }

void refdom_destroy(void* d) { delete static_cast<Base*>(d); }

// Domain::sync (m == NULL) or Domain::syncGrav; returns 0, or -1 if the reference threw or changed the size
int refdom_sync(void* d, void* x, void* y, void* z, void* h, void* m, uint32_t* ids, size_t n, void* keys_out)
{
    try
    {
        int rc = static_cast<Base*>(d)->sync(x, y, z, h, m, ids, n, keys_out);
        if (rc) lastError = "the reference changed the particle count";
        return rc;
    }
    catch (const std::exception& e)
    {
        lastError = e.what();
        return -1;
    }
    catch (...)
    {
        lastError = "unknown exception";
        return -1;
    }
}

void refdom_sizes(void* d, int64_t* out) { static_cast<Base*>(d)->sizes(out); }
void refdom_box(void* d, double* out) { static_cast<Base*>(d)->box(out); }
void refdom_global(void* d, void* leaves, uint32_t* counts) { static_cast<Base*>(d)->global(leaves, counts); }
void refdom_focus(void* d, void* leaves, uint32_t* leaf_counts, uint32_t* node_counts, uint32_t* layout)
{
    static_cast<Base*>(d)->focus(leaves, leaf_counts, node_counts, layout);
}
void refdom_octree(void* d, void* prefixes, int32_t* child_offsets, int32_t* internal_to_leaf, int32_t* level_range,
                   void* centers, void* sizes)
{
    static_cast<Base*>(d)->octree(prefixes, child_offsets, internal_to_leaf, level_range, centers, sizes);
}
void refdom_expansion_centers(void* d, void* centers) { static_cast<Base*>(d)->expansion(centers); }

} // extern "C"
