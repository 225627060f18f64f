// a16: radius neighbor search with the per-particle semantics of the reference CPU path.
//
// Reference semantics (findneighbors.hpp:77-177, traversal.hpp:26-69, boxoverlap.hpp:217-250, util/array.hpp:237-240):
// particle i visits node n iff every node on the root->n path passes
//     norm2(minDistance(p_i, center_n, size_n)) < 4 h_i^2 * ext^2      (right-fold norm2, CPU minDistance order)
// and at each visited leaf records j != i with  dx*dx + dy*dy + dz*dz < 4 h_i^2  (left fold), ascending j,
// keeping the first ngmax and counting all.
//
// GPU design (wave-cooperative stackless walk): one wave64 owns 64 consecutive targets (compact in space because
// they are SFC-ordered) and walks the tree once, with a wave-uniform node cursor driven by childOffsets/parents.
// Each lane carries `dead`, the depth at which its own predicate chain last failed, so at every node a lane
// evaluates its own test only if all its ancestors passed — exactly the CPU's visited set, not the superset a
// group-box walk would give. The wave descends where any lane passes. At a leaf, coordinates of candidate j are
// wave-uniform (scalar loads) and every lane tests them against its own particle, which also keeps the per-lane
// neighbor order ascending. A step cap (2 * numNodes + 8) bounds the walk even for a malformed tree.
#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

constexpr int kNbThreads = 256;
constexpr int kDeadNever = 1 << 30;

template<class T, bool AnyPbc>
__global__ __launch_bounds__(kNbThreads) void findNeighborsKernel(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, const T* __restrict__ h,
    uint32_t firstId, uint32_t lastId, typename BoxOf<T>::type box, const int32_t* __restrict__ childOffsets,
    const int32_t* __restrict__ parents, const int32_t* __restrict__ internalToLeaf,
    const uint32_t* __restrict__ layout, const T* __restrict__ centers, const T* __restrict__ sizes, int32_t numNodes,
    float ext, uint32_t ngmax, uint32_t* __restrict__ neighbors, uint32_t* __restrict__ neighborsCount)
{
    const int lane        = laneId();
    const uint32_t group  = uniform(uint32_t((size_t(blockIdx.x) * kNbThreads + threadIdx.x) / kWave));
    const uint64_t first  = uint64_t(firstId) + uint64_t(group) * kWave;
    if (first >= lastId) return;
    const uint32_t i = uint32_t(first) + uint32_t(lane);
    const bool valid = i < lastId;

    T xi = 0, yi = 0, zi = 0, hi = 0;
    if (valid)
    {
        xi = x[i];
        yi = y[i];
        zi = z[i];
        hi = h[i];
    }
    const T radiusSq = T(4.0) * hi * hi;
    const T cellR2   = radiusSq * T(ext) * T(ext);

    // periodic images only for particles whose 2h box leaves the domain (findneighbors.hpp:102-104)
    bool px = false, py = false, pz = false;
    if constexpr (AnyPbc)
    {
        T s2        = T(2) * hi;
        bool inside = (xi - s2) >= box.lim[0] && (yi - s2) >= box.lim[2] && (zi - s2) >= box.lim[4] &&
                      (xi + s2) <= box.lim[1] && (yi + s2) <= box.lim[3] && (zi + s2) <= box.lim[5];
        px          = !inside && box.bnd[0] == HIPCSTONE_PERIODIC;
        py          = !inside && box.bnd[1] == HIPCSTONE_PERIODIC;
        pz          = !inside && box.bnd[2] == HIPCSTONE_PERIODIC;
    }
    auto fold = [&](T d, bool p, int dim) -> T {
        if constexpr (AnyPbc) { return p ? d - box.len[dim] * rint(d * box.ilen[dim]) : d; }
        else { return d; }
    };

    auto nodePasses = [&](int32_t node) -> bool {
        const T* c = centers + size_t(node) * 3;
        const T* s = sizes + size_t(node) * 3;
        T d0       = fabs(fold(c[0] - xi, px, 0)) - s[0];
        T d1       = fabs(fold(c[1] - yi, py, 1)) - s[1];
        T d2       = fabs(fold(c[2] - zi, pz, 2)) - s[2];
        d0 += fabs(d0);
        d1 += fabs(d1);
        d2 += fabs(d2);
        d0 *= T(0.5);
        d1 *= T(0.5);
        d2 *= T(0.5);
        return (d0 * d0 + (d1 * d1 + d2 * d2)) < cellR2;
    };

    uint32_t count = 0;
    uint32_t* row  = neighbors ? neighbors + size_t(i - firstId) * ngmax : nullptr;

    auto leafAction = [&](int32_t node, bool act) {
        int32_t leafIdx = uniform(internalToLeaf[node]);
        uint32_t jb     = uniform(layout[leafIdx]);
        uint32_t je     = uniform(layout[leafIdx + 1]);
        for (uint32_t j = jb; j < je; ++j)
        {
            T xx = fold(x[j] - xi, px, 0);
            T yy = fold(y[j] - yi, py, 1);
            T zz = fold(z[j] - zi, pz, 2);
            bool hit = act && j != i && (xx * xx + yy * yy + zz * zz < radiusSq);
            if (hit)
            {
                if (count < ngmax) row[count] = j;
                count++;
            }
        }
    };

    int dead = valid ? kDeadNever : -1;

    // root (traversal.hpp:35-45)
    bool pass = valid && nodePasses(0);
    if (ballot(pass) != 0)
    {
        int32_t rootChild = uniform(childOffsets[0]);
        if (rootChild == 0) { leafAction(0, pass); }
        else
        {
            dead             = pass ? kDeadNever : (valid ? 0 : -1);
            int32_t node     = rootChild;
            int depth        = 1;
            bool backtrack   = false;
            int64_t stepsLeft = 2 * int64_t(numNodes) + 8;
            while (node != 0 && stepsLeft-- > 0)
            {
                if (node < 0 || node >= numNodes) break;
                int32_t co  = uniform(childOffsets[node]);
                bool isLeaf = co == 0;
                bool descend = false;
                if (!backtrack)
                {
                    bool eligible = dead >= depth;
                    bool t        = eligible && nodePasses(node);
                    if (eligible) dead = t ? kDeadNever : depth;
                    descend = ballot(t) != 0;
                    if (isLeaf && descend) leafAction(node, t);
                }
                int32_t sib = (node - 1) % 8;
                if (!isLeaf && descend)
                {
                    node      = co;
                    depth     = depth + 1;
                    backtrack = false;
                }
                else if (sib < 7)
                {
                    node      = node + 1;
                    backtrack = false;
                }
                else
                {
                    node      = uniform(parents[(node - 1) / 8]);
                    depth     = depth - 1;
                    backtrack = true;
                }
            }
        }
    }
    if (valid) neighborsCount[i - firstId] = count;
}

template<class T>
static hipcstone_status findNeighbors(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h,
                                      uint32_t firstId, uint32_t lastId, const typename BoxOf<T>::type* box,
                                      const int32_t* childOffsets, const int32_t* parents,
                                      const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                      const T* sizes, int32_t numNodes, float ext, uint32_t ngmax,
                                      uint32_t* neighbors, uint32_t* neighborsCount)
{
    if (lastId <= firstId) return HIPCSTONE_OK;
    if (!x || !y || !z || !h || !box || !childOffsets || !parents || !internalToLeaf || !layout || !centers ||
        !sizes || !neighborsCount)
        return invalid("findNeighbors: null pointer");
    if (ngmax > 0 && !neighbors) return invalid("findNeighbors: neighbors is null with ngmax > 0");
    if (numNodes <= 0) return invalid("findNeighbors: numNodes must be positive");
    size_t groups   = (size_t(lastId - firstId) + kWave - 1) / kWave;
    size_t threads  = groups * kWave;
    unsigned blocks = ceilDiv(threads, kNbThreads);
    bool anyPbc     = box->bnd[0] == HIPCSTONE_PERIODIC || box->bnd[1] == HIPCSTONE_PERIODIC ||
                  box->bnd[2] == HIPCSTONE_PERIODIC;
    if (anyPbc)
        findNeighborsKernel<T, true><<<blocks, kNbThreads, 0, stream(s)>>>(
            x, y, z, h, firstId, lastId, *box, childOffsets, parents, internalToLeaf, layout, centers, sizes, numNodes,
            ext, ngmax, neighbors, neighborsCount);
    else
        findNeighborsKernel<T, false><<<blocks, kNbThreads, 0, stream(s)>>>(
            x, y, z, h, firstId, lastId, *box, childOffsets, parents, internalToLeaf, layout, centers, sizes, numNodes,
            ext, ngmax, neighbors, neighborsCount);
    HC_CHECK_LAUNCH("findNeighborsKernel");
    return HIPCSTONE_OK;
}

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
                                layout, centers, sizes, num_nodes, ext, ngmax, neighbors, neighbors_count);            \
    }

HIPCSTONE_DEF_NEIGHBORS(f32, float)
HIPCSTONE_DEF_NEIGHBORS(f64, double)

} // extern "C"
