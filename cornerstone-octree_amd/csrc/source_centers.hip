// Focus-tree source (mass) centers and MAC radii: the expansion centers a gravity client reads after syncGrav.
//
// Reference semantics (paths under include/cstone/): focus/source_center.hpp:48-78 (addBody, normalizeMass,
// massCenter), :110-125 (computeLeafMassCenter, the CPU path), :82-98 (CombineSourceCenter) with tree/octree.hpp:572-590
// (upsweep), :127-140 (setMac), :160-168 (geoMacSpheres); traversal/macs.hpp computeVecMacR2 / computeMinMacR2 with
// sfc/box.hpp:318-335 centerAndSize of the node's Hilbert box (sfcKey is the Hilbert curve, sfc/sfc.hpp:38-40).
//
// A leaf's mass center sums its particles in index order, one thread per leaf, as the reference CPU path does (the
// reference GPU kernel splits a leaf over several lanes and reduces by shuffles, a different rounding order); every
// product and sum is rounded on its own (no contraction), so the results are bit-identical to the CPU path. Centers
// are 4 values per node: the mass center x, y, z and the total |m|, which setMac replaces by the squared MAC radius.
#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

template<class T>
struct alignas(4 * sizeof(T)) Center4
{
    T x, y, z, m;
};

template<class T>
__device__ inline void addBody(Center4<T>& c, T x, T y, T z, T m)
{
    T w = fabs(m);
    c.x += w * x;
    c.y += w * y;
    c.z += w * z;
    c.m += w;
}

template<class T>
__device__ inline Center4<T> normalizeMass(Center4<T> c)
{
    T invM = (c.m != T(0)) ? T(1) / c.m : T(1);
    c.x *= invM;
    c.y *= invM;
    c.z *= invM;
    return c;
}

// one block per 256 consecutive leaves; their particles are one contiguous range. If it fits, the block stages it
// in LDS with coalesced loads first (each thread's own leaf is then summed from LDS, still in index order); a block
// whose range is larger reads its particles from global memory directly
constexpr int kMcLeaves     = 256;
constexpr int kMcStageBytes = 32768; // LDS staging per block: 2048 fp32 or 1024 fp64 particles

template<class T>
__global__ __launch_bounds__(kMcLeaves) void leafMassCentersKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                                   const T* __restrict__ z, const T* __restrict__ m,
                                                                   const int32_t* __restrict__ leafToInternal,
                                                                   int32_t numLeaves,
                                                                   const uint32_t* __restrict__ layout,
                                                                   Center4<T>* __restrict__ centers)
{
    constexpr uint32_t kStage = kMcStageBytes / sizeof(Center4<T>);
    __shared__ Center4<T> stage[kStage];
    const int32_t first = int32_t(blockIdx.x) * kMcLeaves;
    const int32_t last  = first + kMcLeaves < numLeaves ? first + kMcLeaves : numLeaves;
    const int32_t leaf  = first + int32_t(threadIdx.x);
    const uint32_t base = layout[first];
    const uint32_t span = layout[last] - base;
    const bool staged   = span <= kStage;
    if (staged)
    {
        for (uint32_t i = threadIdx.x; i < span; i += kMcLeaves)
        {
            const uint32_t g = base + i;
            stage[i]         = Center4<T>{x[g], y[g], z[g], m[g]};
        }
    }
    __syncthreads();
    if (leaf >= numLeaves) return;
    Center4<T> c{0, 0, 0, 0};
    const uint32_t b = layout[leaf], e = layout[leaf + 1];
    if (staged)
    {
        for (uint32_t i = b - base; i < e - base; ++i)
        {
            const Center4<T> p = stage[i];
            addBody(c, p.x, p.y, p.z, p.m);
        }
    }
    else
    {
        for (uint32_t i = b; i < e; ++i)
        {
            addBody(c, x[i], y[i], z[i], m[i]);
        }
    }
    centers[leafToInternal[leaf]] = normalizeMass(c);
}

// one tree level of the upsweep: every internal node in [first, last) becomes the mass center of its 8 children
template<class T>
__global__ __launch_bounds__(256) void upsweepCentersKernel(int32_t first, int32_t last,
                                                            const int32_t* __restrict__ childOffsets,
                                                            Center4<T>* __restrict__ centers)
{
    int32_t i = first + int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= last) return;
    int32_t child = childOffsets[i];
    if (child == 0) return;
    Center4<T> c{0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; ++k)
    {
        Center4<T> s = centers[child + k];
        addBody(c, s.x, s.y, s.z, s.m);
    }
    centers[i] = normalizeMass(c);
}

// geometric center and half size of a node from its Hilbert box (nodeFpCenters / centerAndSize)
template<class K, class T>
__device__ inline void geoCenterSize(K prefix, const typename BoxOf<T>::type& box, T* gc, T* gs)
{
    K startKey     = decodePlaceholderBit(prefix);
    unsigned level = decodePrefixLength(prefix) / 3;
    int ib[6];
    hilbertIBox(startKey, level, ib);
    constexpr T uL = T(1.) / T(1u << maxLevel<K>());
    T hx           = T(0.5) * uL * box.len[0];
    T hy           = T(0.5) * uL * box.len[1];
    T hz           = T(0.5) * uL * box.len[2];
    gc[0]          = box.lim[0] + T(ib[1] + ib[0]) * hx;
    gc[1]          = box.lim[2] + T(ib[3] + ib[2]) * hy;
    gc[2]          = box.lim[4] + T(ib[5] + ib[4]) * hz;
    gs[0]          = T(ib[1] - ib[0]) * hx;
    gs[1]          = T(ib[3] - ib[2]) * hy;
    gs[2]          = T(ib[5] - ib[4]) * hz;
}

// util::max of a Vec3, left to right
template<class T>
__device__ inline T maxOf3(const T* v)
{
    T r = v[0];
    r   = r > v[1] ? r : v[1];
    r   = r > v[2] ? r : v[2];
    return r;
}

// setMac: centers[i].m = computeVecMacR2(prefix, center, invTheta) = (l * invTheta + |center - geoCenter|)^2 with
// l = 2 * max(geoSize), when the node's mass is nonzero, else 0
template<class K, class T>
__global__ __launch_bounds__(256) void setMacKernel(const K* __restrict__ prefixes, int32_t numNodes,
                                                    Center4<T>* __restrict__ centers, float invTheta,
                                                    typename BoxOf<T>::type box)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    Center4<T> c = centers[i];
    T gc[3], gs[3];
    geoCenterSize<K, T>(prefixes[i], box, gc, gs);
    T dx        = c.x - gc[0];
    T dy        = c.y - gc[1];
    T dz        = c.z - gc[2];
    T s         = sqrt(dx * dx + (dy * dy + dz * dz));
    T l         = T(2.0) * maxOf3(gs);
    T mac       = l * T(invTheta) + s;
    centers[i].m = (c.m != T(0)) ? mac * mac : T(0);
}

// geoMacSpheres: centers[i] = {geoCenter, (l * invTheta)^2} (computeMinMacR2)
template<class K, class T>
__global__ __launch_bounds__(256) void geoMacSpheresKernel(const K* __restrict__ prefixes, int32_t numNodes,
                                                           Center4<T>* __restrict__ centers, float invTheta,
                                                           typename BoxOf<T>::type box)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    T gc[3], gs[3];
    geoCenterSize<K, T>(prefixes[i], box, gc, gs);
    T l        = T(2) * maxOf3(gs);
    T mac      = l * T(invTheta);
    centers[i] = Center4<T>{gc[0], gc[1], gc[2], mac * mac};
}

template<class T>
static hipcstone_status leafMassCenters(hipcstone_stream s, const T* x, const T* y, const T* z, const T* m,
                                        const int32_t* leafToInternal, int32_t numLeaves, const uint32_t* layout,
                                        T* centers)
{
    if (numLeaves <= 0) return HIPCSTONE_OK;
    if (!x || !y || !z || !m || !leafToInternal || !layout || !centers)
        return invalid("leafMassCenters: null pointer");
    leafMassCentersKernel<T><<<ceilDiv(size_t(numLeaves), kMcLeaves), kMcLeaves, 0, stream(s)>>>(
        x, y, z, m, leafToInternal, numLeaves, layout, reinterpret_cast<Center4<T>*>(centers));
    HC_CHECK_LAUNCH("leafMassCentersKernel");
    return HIPCSTONE_OK;
}

template<class T>
static hipcstone_status upsweepCenters(hipcstone_stream s, int32_t numLevels, const int32_t* levelRangeHost,
                                       const int32_t* childOffsets, T* centers)
{
    if (numLevels < 0 || numLevels > 32) return invalid("upsweepCenters: numLevels outside [0, 32]");
    if (!levelRangeHost || !childOffsets || !centers) return invalid("upsweepCenters: null pointer");
    for (int level = numLevels; level >= 0; --level)
    {
        int32_t first = levelRangeHost[level], last = levelRangeHost[level + 1];
        if (first < 0 || last < first) return invalid("upsweepCenters: level_range is not ascending");
        if (last == first) continue;
        upsweepCentersKernel<T><<<ceilDiv(size_t(last - first), 256), 256, 0, stream(s)>>>(
            first, last, childOffsets, reinterpret_cast<Center4<T>*>(centers));
        HC_CHECK_LAUNCH("upsweepCentersKernel");
    }
    return HIPCSTONE_OK;
}

template<class K, class T>
static hipcstone_status setMac(hipcstone_stream s, const K* prefixes, int32_t numNodes, T* centers, float invTheta,
                               const typename BoxOf<T>::type* box)
{
    if (numNodes <= 0) return HIPCSTONE_OK;
    if (!prefixes || !centers || !box) return invalid("setMac: null pointer");
    setMacKernel<K, T><<<ceilDiv(size_t(numNodes), 256), 256, 0, stream(s)>>>(
        prefixes, numNodes, reinterpret_cast<Center4<T>*>(centers), invTheta, *box);
    HC_CHECK_LAUNCH("setMacKernel");
    return HIPCSTONE_OK;
}

template<class K, class T>
static hipcstone_status geoMacSpheres(hipcstone_stream s, const K* prefixes, int32_t numNodes, T* centers,
                                      float invTheta, const typename BoxOf<T>::type* box)
{
    if (numNodes <= 0) return HIPCSTONE_OK;
    if (!prefixes || !centers || !box) return invalid("geoMacSpheres: null pointer");
    geoMacSpheresKernel<K, T><<<ceilDiv(size_t(numNodes), 256), 256, 0, stream(s)>>>(
        prefixes, numNodes, reinterpret_cast<Center4<T>*>(centers), invTheta, *box);
    HC_CHECK_LAUNCH("geoMacSpheresKernel");
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

#define HIPCSTONE_DEF_SOURCE_CENTERS(RS, T)                                                                            \
    hipcstone_status hipcstone_leaf_mass_centers_##RS(hipcstone_stream s, const T* x, const T* y, const T* z,         \
                                                      const T* m, const int32_t* leaf_to_internal,                    \
                                                      int32_t num_leaves, const uint32_t* layout, T* centers)         \
    {                                                                                                                  \
        return leafMassCenters<T>(s, x, y, z, m, leaf_to_internal, num_leaves, layout, centers);                      \
    }                                                                                                                  \
    hipcstone_status hipcstone_upsweep_centers_##RS(hipcstone_stream s, int32_t num_levels,                          \
                                                    const int32_t* level_range_host, const int32_t* child_offsets,   \
                                                    T* centers)                                                        \
    {                                                                                                                  \
        return upsweepCenters<T>(s, num_levels, level_range_host, child_offsets, centers);                            \
    }

#define HIPCSTONE_DEF_MACS(KS, K, RS, T)                                                                               \
    hipcstone_status hipcstone_set_mac_##KS##_##RS(hipcstone_stream s, const K* prefixes, int32_t num_nodes,          \
                                                   T* centers, float inv_theta, const hipcstone_box_##RS* box)        \
    {                                                                                                                  \
        return setMac<K, T>(s, prefixes, num_nodes, centers, inv_theta, box);                                          \
    }                                                                                                                  \
    hipcstone_status hipcstone_geo_mac_spheres_##KS##_##RS(hipcstone_stream s, const K* prefixes, int32_t num_nodes,  \
                                                           T* centers, float inv_theta,                               \
                                                           const hipcstone_box_##RS* box)                             \
    {                                                                                                                  \
        return geoMacSpheres<K, T>(s, prefixes, num_nodes, centers, inv_theta, box);                                   \
    }

HIPCSTONE_DEF_SOURCE_CENTERS(f32, float)
HIPCSTONE_DEF_SOURCE_CENTERS(f64, double)
HIPCSTONE_DEF_MACS(u32, uint32_t, f32, float)
HIPCSTONE_DEF_MACS(u32, uint32_t, f64, double)
HIPCSTONE_DEF_MACS(u64, uint64_t, f32, float)
HIPCSTONE_DEF_MACS(u64, uint64_t, f64, double)

} // extern "C"
