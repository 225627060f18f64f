// a15 and its callers: SFC assignment, halo discovery driver, the local (assigned + halo) particle layout, marked
// leaf ranges and the range gather that packs halo particles into local arrays.
//
// Reference semantics: domain/domaindecomp.hpp:32-55 (uniformBins) and :103-118 (makeSfcAssignment),
// focus/octree_focus_mpi.hpp:511-568 (discoverHalos), domain/layout.hpp:111-141 (extractMarkedElements) and
// :153-185 (computeNodeLayout), halos/gather_halos_gpu.cu:28-57 (gatherRanges).
#include "common.hpp"
#include "runtime.hpp"
#include "scan.hpp"

namespace hipcstone
{

// ---- uniformBins / makeSfcAssignment -------------------------------------------------------------------------------
// countScan[0..numLeaves] is the uint64 exclusive scan of the leaf counts with the total in countScan[numLeaves]
// (the reference's inclusive scan with a leading 0). Bin i in (0, numBins) starts at the first leaf whose scan
// reaches uint64(i * (total / numBins)).
template<class K>
__global__ void uniformBinsKernel(const uint64_t* __restrict__ countScan, int32_t numLeaves, int32_t numBins,
                                  const K* __restrict__ leaves, int32_t* __restrict__ treeOffsets,
                                  K* __restrict__ boundaries)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > numBins) return;
    int32_t bin;
    if (i == 0) bin = 0;
    else if (i == numBins) bin = numLeaves;
    else
    {
        double binCount = double(countScan[numLeaves]) / double(numBins);
        uint64_t target = uint64_t(double(i) * binCount);
        // lower_bound over countScan[0 .. numLeaves]
        int32_t lo = 0, len = numLeaves + 1;
        while (len > 0)
        {
            int32_t half = len >> 1;
            if (countScan[lo + half] < target)
            {
                lo = lo + half + 1;
                len -= half + 1;
            }
            else { len = half; }
        }
        bin = lo;
    }
    treeOffsets[i] = bin;
    boundaries[i]  = leaves[bin];
}

// the reference stores per-rank counts as LocalIndex (uint32), so sums of 2^32 or more wrap as they do there
__global__ void binCountsKernel(const uint64_t* __restrict__ countScan, const int32_t* __restrict__ treeOffsets,
                                int32_t numBins, uint32_t* __restrict__ binCounts)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= numBins) return;
    binCounts[i] = uint32_t(countScan[treeOffsets[i + 1]] - countScan[treeOffsets[i]]);
}

static size_t assignmentWorkspace(int32_t numLeaves)
{
    size_t L = numLeaves > 0 ? size_t(numLeaves) : 0;
    return (((L + 1) * sizeof(uint64_t) + 255) & ~size_t(255)) + scan::workspaceBytes<uint64_t>(L);
}

template<class K>
static hipcstone_status sfcAssignment(hipcstone_stream s, const K* leaves, const uint32_t* counts, int32_t numLeaves,
                                      int32_t numRanks, int32_t* treeOffsets, K* boundaries, uint32_t* rankCounts,
                                      void* ws, size_t wsBytes)
{
    if (numLeaves <= 0 || numRanks <= 0) return invalid("makeSfcAssignment: need leaves and ranks");
    if (!leaves || !counts || !treeOffsets || !boundaries || !rankCounts)
        return invalid("makeSfcAssignment: null pointer");
    if (!ws || wsBytes < assignmentWorkspace(numLeaves)) return invalid("makeSfcAssignment: workspace too small");
    uint64_t* countScan = static_cast<uint64_t*>(ws);
    void* scanWs = static_cast<uint8_t*>(ws) + (((size_t(numLeaves) + 1) * sizeof(uint64_t) + 255) & ~size_t(255));
    if (auto e = scan::exclusiveScan<uint64_t, uint32_t>(stream(s), counts, countScan, size_t(numLeaves), 0ull,
                                                         countScan + numLeaves, scanWs,
                                                         scan::workspaceBytes<uint64_t>(size_t(numLeaves))))
        return e;
    unsigned blocks = unsigned(ceilDiv(size_t(numRanks) + 1, 256));
    uniformBinsKernel<K><<<blocks, 256, 0, stream(s)>>>(countScan, numLeaves, numRanks, leaves, treeOffsets,
                                                        boundaries);
    HC_CHECK_LAUNCH("uniformBinsKernel");
    binCountsKernel<<<blocks, 256, 0, stream(s)>>>(countScan, treeOffsets, numRanks, rankCounts);
    HC_CHECK_LAUNCH("binCountsKernel");
    return HIPCSTONE_OK;
}

// ---- discoverHalos: search boxes seeded with the leaves' geometric centers -----------------------------------------
template<class T>
__global__ void gatherLeafCentersKernel(const int32_t* __restrict__ leafToInternal, const T* __restrict__ nodeCenters,
                                        int32_t first, int32_t last, T* __restrict__ searchCenters)
{
    int32_t leaf = first + int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (leaf >= last) return;
    size_t src = size_t(leafToInternal[leaf]) * 3, dst = size_t(leaf) * 3;
    searchCenters[dst + 0] = nodeCenters[src + 0];
    searchCenters[dst + 1] = nodeCenters[src + 1];
    searchCenters[dst + 2] = nodeCenters[src + 2];
}

__global__ void zeroBytesKernel(uint8_t* __restrict__ p, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---- computeNodeLayout ---------------------------------------------------------------------------------------------
// A leaf keeps its particles iff it is assigned (first <= i < last) or its internal node index is flagged as a halo;
// layout = exclusive scan of the kept counts with the total in layout[numLeaves]. This kernel writes the kept
// counts, the scan runs in place afterwards.
__global__ void keptCountsKernel(const uint32_t* __restrict__ counts, const uint8_t* __restrict__ flags,
                                 const int32_t* __restrict__ leafToInternal, int32_t numLeaves, int32_t first,
                                 int32_t last, uint32_t* __restrict__ kept)
{
    int32_t i = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= numLeaves) return;
    bool have = (first <= i && i < last) || flags[leafToInternal[i]] != 0;
    kept[i]   = have ? counts[i] : 0u;
}

// ---- extractMarkedElements -----------------------------------------------------------------------------------------
// Leaf i in [first, second) is marked iff flags[i+1] > flags[i] (flags is a layout, so nondecreasing). Each maximal
// run [a, b) of marked leaves yields the pair (source[a], source[b]), in order. starts[k] = 1 where a run begins at
// first + k; the run index of a start is the exclusive scan of starts.
__global__ void runStartsKernel(const uint32_t* __restrict__ flags, int32_t first, int32_t second,
                                uint32_t* __restrict__ starts)
{
    int32_t k = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    int32_t i = first + k;
    if (i >= second) return;
    bool marked = flags[i + 1] > flags[i];
    bool prev   = i > first && flags[i] > flags[i - 1];
    starts[k]   = (marked && !prev) ? 1u : 0u;
}

// runIndex holds the exclusive scan of starts for [first, second) plus the run total at [second - first].
template<class V>
__global__ void runPairsKernel(const V* __restrict__ source, const uint32_t* __restrict__ flags,
                               const uint32_t* __restrict__ runIndex, int32_t first, int32_t second,
                               V* __restrict__ pairs, uint32_t maxPairs, uint32_t* __restrict__ numPairs)
{
    int32_t k = int32_t(blockIdx.x * blockDim.x + threadIdx.x);
    int32_t i = first + k;
    if (i > second) return;
    uint32_t total = runIndex[second - first];
    if (k == 0) numPairs[0] = total;
    if (i < second)
    {
        bool marked = flags[i + 1] > flags[i];
        bool prev   = i > first && flags[i] > flags[i - 1];
        if (marked && !prev)
        {
            uint32_t r = runIndex[k];
            if (r < maxPairs) pairs[2 * size_t(r)] = source[i];
        }
    }
    // a run ends at i when leaf i-1 is marked and leaf i is not, or at i == second after a marked leaf
    if (i > first)
    {
        bool prevMarked = flags[i] > flags[i - 1];
        bool marked     = i < second && flags[i + 1] > flags[i];
        if (prevMarked && !marked)
        {
            uint32_t r = runIndex[k] - 1; // runs that started before i, minus one
            if (r < maxPairs) pairs[2 * size_t(r) + 1] = source[i];
        }
    }
}

// ---- gatherRanges ---------------------------------------------------------------------------------------------------
// buffer[t] = src[rangeOffsets[r] + t - rangeScan[r]] with r = upper_bound(rangeScan, t) - 1. Each block owns a
// contiguous chunk of the buffer; the ranges that can cover the chunk start at r0 (one search per block) and are
// staged in LDS, so the per-element search runs in LDS. Elements are moved as opaque records of E bytes.
template<int E>
struct Record
{
    uint8_t b[E];
};
template<>
struct alignas(4) Record<4>
{
    uint32_t v;
};
template<>
struct alignas(8) Record<8>
{
    uint64_t v;
};
template<>
struct alignas(4) Record<12>
{
    uint32_t v[3];
};
template<>
struct alignas(16) Record<16>
{
    uint64_t v[2];
};

constexpr int kGrThreads = 256;
constexpr int kGrItems   = 8;
constexpr int kGrChunk   = kGrThreads * kGrItems;
constexpr int kGrWindow  = 512;

__device__ inline uint32_t upperBoundU32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t v)
{
    uint32_t len = hi - lo;
    while (len > 0)
    {
        uint32_t half = len >> 1;
        if (a[lo + half] <= v)
        {
            lo += half + 1;
            len -= half + 1;
        }
        else { len = half; }
    }
    return lo;
}

template<int E>
__global__ __launch_bounds__(kGrThreads) void gatherRangesKernel(const uint32_t* __restrict__ rangeScan,
                                                                 const uint32_t* __restrict__ rangeOffsets,
                                                                 uint32_t numRanges, const Record<E>* __restrict__ src,
                                                                 Record<E>* __restrict__ buffer, size_t bufferSize)
{
    __shared__ uint32_t win[kGrWindow];
    __shared__ uint32_t r0Sh;
    size_t chunk = size_t(blockIdx.x) * kGrChunk;
    if (threadIdx.x == 0)
    {
        uint32_t ub = upperBoundU32(rangeScan, 0, numRanges, uint32_t(chunk));
        r0Sh        = ub > 0 ? ub - 1 : 0;
    }
    __syncthreads();
    uint32_t r0 = r0Sh;
    for (int k = threadIdx.x; k < kGrWindow; k += kGrThreads)
        win[k] = r0 + k < numRanges ? rangeScan[r0 + k] : 0xFFFFFFFFu;
    __syncthreads();
    size_t chunkEnd = chunk + kGrChunk < bufferSize ? chunk + kGrChunk : bufferSize;
    // every element's range lies inside the window iff the window ends past the chunk (or holds all ranges)
    bool inWindow = r0 + kGrWindow >= numRanges || size_t(win[kGrWindow - 1]) >= chunkEnd;
#pragma unroll
    for (int it = 0; it < kGrItems; ++it)
    {
        size_t t = chunk + size_t(it) * kGrThreads + threadIdx.x;
        if (t >= bufferSize) break;
        uint32_t ub;
        if (inWindow)
        {
            uint32_t hi = numRanges - r0 < uint32_t(kGrWindow) ? numRanges - r0 : uint32_t(kGrWindow);
            ub          = r0 + upperBoundU32(win, 0, hi, uint32_t(t));
        }
        else { ub = upperBoundU32(rangeScan, r0, numRanges, uint32_t(t)); }
        if (ub == 0 || rangeScan[ub - 1] > t) continue; // no range starts at or before t: nothing to copy
        uint32_t r = ub - 1;
        buffer[t]  = src[size_t(rangeOffsets[r]) + (t - rangeScan[r])];
    }
}

// ---- fused halo gather: one wave per leaf, particles move from the source layout to the local layout -----------------
template<class T>
__global__ __launch_bounds__(256) void gatherLeafRangesKernel(const uint32_t* __restrict__ srcLayout,
                                                              const uint32_t* __restrict__ dstLayout,
                                                              int32_t numLeaves, const T* __restrict__ s0,
                                                              const T* __restrict__ s1, const T* __restrict__ s2,
                                                              const T* __restrict__ s3, T* __restrict__ d0,
                                                              T* __restrict__ d1, T* __restrict__ d2,
                                                              T* __restrict__ d3)
{
    int32_t leaf = int32_t((size_t(blockIdx.x) * blockDim.x + threadIdx.x) / kWave);
    if (leaf >= numLeaves) return;
    const uint32_t lane = uint32_t(laneId());
    uint32_t db = dstLayout[leaf], de = dstLayout[leaf + 1];
    uint32_t sb = srcLayout[leaf];
    for (uint32_t k = lane; db + k < de; k += kWave)
    {
        if (s0) d0[db + k] = s0[sb + k];
        if (s1) d1[db + k] = s1[sb + k];
        if (s2) d2[db + k] = s2[sb + k];
        if (s3) d3[db + k] = s3[sb + k];
    }
}

// ---- host side -----------------------------------------------------------------------------------------------------
static hipcstone_status nodeLayout(hipcstone_stream s, const uint32_t* counts, const uint8_t* flags,
                                   const int32_t* leafToInternal, int32_t numLeaves, int32_t first, int32_t last,
                                   uint32_t* layout, void* ws, size_t wsBytes)
{
    if (numLeaves <= 0) return invalid("computeNodeLayout: numLeaves must be positive");
    if (!counts || !flags || !leafToInternal || !layout) return invalid("computeNodeLayout: null pointer");
    if (first < 0 || last < first || last > numLeaves) return invalid("computeNodeLayout: bad assigned range");
    keptCountsKernel<<<unsigned(ceilDiv(size_t(numLeaves), 256)), 256, 0, stream(s)>>>(
        counts, flags, leafToInternal, numLeaves, first, last, layout);
    HC_CHECK_LAUNCH("keptCountsKernel");
    return scan::exclusiveScan<uint32_t>(stream(s), layout, layout, size_t(numLeaves), 0u, layout + numLeaves, ws,
                                         wsBytes);
}

static size_t extractWorkspace(int32_t numElements)
{
    size_t n = numElements > 0 ? size_t(numElements) : 0;
    return (((n + 1) * sizeof(uint32_t) + 255) & ~size_t(255)) + scan::workspaceBytes(n);
}

template<class V>
static hipcstone_status extractMarked(hipcstone_stream s, const V* source, const uint32_t* flags, int32_t first,
                                      int32_t second, V* pairs, uint32_t maxPairs, uint32_t* numPairs, void* ws,
                                      size_t wsBytes)
{
    if (first < 0 || second < first) return invalid("extractMarkedElements: bad range");
    if (!source || !flags || !numPairs || (maxPairs && !pairs)) return invalid("extractMarkedElements: null pointer");
    int32_t n = second - first;
    if (!ws || wsBytes < extractWorkspace(n)) return invalid("extractMarkedElements: workspace too small");
    uint32_t* runIndex = static_cast<uint32_t*>(ws);
    void* scanWs       = static_cast<uint8_t*>(ws) + (((size_t(n) + 1) * sizeof(uint32_t) + 255) & ~size_t(255));
    if (n > 0)
    {
        runStartsKernel<<<unsigned(ceilDiv(size_t(n), 256)), 256, 0, stream(s)>>>(flags, first, second, runIndex);
        HC_CHECK_LAUNCH("runStartsKernel");
    }
    if (auto e = scan::exclusiveScan<uint32_t>(stream(s), runIndex, runIndex, size_t(n), 0u, runIndex + n, scanWs,
                                               wsBytes - (static_cast<uint8_t*>(scanWs) - static_cast<uint8_t*>(ws))))
        return e;
    runPairsKernel<V><<<unsigned(ceilDiv(size_t(n) + 1, 256)), 256, 0, stream(s)>>>(source, flags, runIndex, first,
                                                                                    second, pairs, maxPairs, numPairs);
    HC_CHECK_LAUNCH("runPairsKernel");
    return HIPCSTONE_OK;
}

template<int E>
static hipcstone_status gatherRangesE(hipcstone_stream s, const uint32_t* rangeScan, const uint32_t* rangeOffsets,
                                      int32_t numRanges, const void* src, void* buffer, size_t bufferSize)
{
    size_t blocks = ceilDiv(bufferSize, size_t(kGrChunk));
    gatherRangesKernel<E><<<unsigned(blocks), kGrThreads, 0, stream(s)>>>(
        rangeScan, rangeOffsets, uint32_t(numRanges), static_cast<const Record<E>*>(src),
        static_cast<Record<E>*>(buffer), bufferSize);
    HC_CHECK_LAUNCH("gatherRangesKernel");
    return HIPCSTONE_OK;
}

static hipcstone_status gatherRanges(hipcstone_stream s, const uint32_t* rangeScan, const uint32_t* rangeOffsets,
                                     int32_t numRanges, const void* src, void* buffer, size_t bufferSize,
                                     size_t elemBytes)
{
    if (bufferSize == 0) return HIPCSTONE_OK;
    if (numRanges <= 0) return invalid("gatherRanges: no ranges for a non-empty buffer");
    if (!rangeScan || !rangeOffsets || !src || !buffer) return invalid("gatherRanges: null pointer");
    if (bufferSize > 0xFFFFFFFFull) return invalid("gatherRanges: buffer larger than 2^32 - 1 elements");
    switch (elemBytes)
    {
        case 1: return gatherRangesE<1>(s, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize);
        case 2: return gatherRangesE<2>(s, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize);
        case 4: return gatherRangesE<4>(s, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize);
        case 8: return gatherRangesE<8>(s, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize);
        case 12: return gatherRangesE<12>(s, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize);
        case 16: return gatherRangesE<16>(s, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize);
        default: return invalid("gatherRanges: element size must be 1, 2, 4, 8, 12 or 16 bytes");
    }
}

template<class T>
static hipcstone_status gatherLeafRanges(hipcstone_stream s, const uint32_t* srcLayout, const uint32_t* dstLayout,
                                         int32_t numLeaves, const T* s0, const T* s1, const T* s2, const T* s3, T* d0,
                                         T* d1, T* d2, T* d3)
{
    if (numLeaves <= 0) return HIPCSTONE_OK;
    if (!srcLayout || !dstLayout) return invalid("gatherLeafRanges: null layout");
    if ((s0 && !d0) || (s1 && !d1) || (s2 && !d2) || (s3 && !d3)) return invalid("gatherLeafRanges: null output");
    size_t threads = size_t(numLeaves) * kWave;
    gatherLeafRangesKernel<T><<<unsigned(ceilDiv(threads, 256)), 256, 0, stream(s)>>>(srcLayout, dstLayout, numLeaves,
                                                                                       s0, s1, s2, s3, d0, d1, d2, d3);
    HC_CHECK_LAUNCH("gatherLeafRangesKernel");
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

size_t hipcstone_sfc_assignment_workspace_size(int32_t num_leaves) { return assignmentWorkspace(num_leaves); }

#define HIPCSTONE_DEF_ASSIGN(KS, K)                                                                                    \
    hipcstone_status hipcstone_make_sfc_assignment_##KS(hipcstone_stream s, const K* leaves, const uint32_t* counts,  \
                                                        int32_t num_leaves, int32_t num_ranks, int32_t* tree_offsets, \
                                                        K* boundaries, uint32_t* rank_counts, void* workspace,         \
                                                        size_t workspace_bytes)                                        \
    {                                                                                                                  \
        return sfcAssignment<K>(s, leaves, counts, num_leaves, num_ranks, tree_offsets, boundaries, rank_counts,       \
                                workspace, workspace_bytes);                                                           \
    }
HIPCSTONE_DEF_ASSIGN(u32, uint32_t)
HIPCSTONE_DEF_ASSIGN(u64, uint64_t)

#define HIPCSTONE_DEF_DISCOVER(KS, K, RS, T)                                                                           \
    hipcstone_status hipcstone_discover_halos_##KS##_##RS(                                                             \
        hipcstone_stream s, const K* prefixes, const int32_t* child_offsets, const int32_t* parents,                  \
        const int32_t* leaf_to_internal, const T* node_centers, const T* node_sizes, const K* leaves,                   \
        int32_t num_nodes, const T* x, const T* y, const T* z, const T* h, const uint32_t* layout,                     \
        int32_t first_node, int32_t last_node, const hipcstone_box_##RS* box, float search_ext, T* search_centers,     \
        T* search_sizes, uint8_t* flags, int32_t accumulate, void* workspace, size_t workspace_bytes)                  \
    {                                                                                                                  \
        if (!leaf_to_internal || !node_centers || !search_centers || !search_sizes || !flags || !box)                 \
            return invalid("discoverHalos: null pointer");                                                            \
        if (num_nodes <= 0 || first_node < 0 || last_node < first_node)                                               \
            return invalid("discoverHalos: bad node range");                                                          \
        if (last_node > first_node)                                                                                    \
        {                                                                                                              \
            gatherLeafCentersKernel<T><<<unsigned(ceilDiv(size_t(last_node - first_node), 256)), 256, 0, stream(s)>>>( \
                leaf_to_internal, node_centers, first_node, last_node, search_centers);                               \
            HC_CHECK_LAUNCH("gatherLeafCentersKernel");                                                                \
            if (auto e = hipcstone_compute_bounding_box_##RS(s, x, y, z, h, layout, first_node, last_node,            \
                                                             T(2.0f * search_ext), search_centers, search_sizes))      \
                return e;                                                                                              \
        }                                                                                                              \
        if (!accumulate)                                                                                               \
        {                                                                                                              \
            zeroBytesKernel<<<unsigned(ceilDiv(size_t(num_nodes), 256)), 256, 0, stream(s)>>>(flags,                \
                                                                                              size_t(num_nodes));      \
            HC_CHECK_LAUNCH("zeroBytesKernel");                                                                        \
        }                                                                                                              \
        return hipcstone_find_halos_##KS##_##RS(s, prefixes, child_offsets, parents, node_centers, node_sizes,        \
                                                leaves, search_centers, search_sizes, box, first_node, last_node,      \
                                                num_nodes, flags, workspace, workspace_bytes);                         \
    }
HIPCSTONE_DEF_DISCOVER(u32, uint32_t, f32, float)
HIPCSTONE_DEF_DISCOVER(u32, uint32_t, f64, double)
HIPCSTONE_DEF_DISCOVER(u64, uint64_t, f32, float)
HIPCSTONE_DEF_DISCOVER(u64, uint64_t, f64, double)

hipcstone_status hipcstone_compute_node_layout(hipcstone_stream s, const uint32_t* counts, const uint8_t* flags,
                                               const int32_t* leaf_to_internal, int32_t num_leaves, int32_t first,
                                               int32_t last, uint32_t* layout, void* workspace,
                                               size_t workspace_bytes)
{
    return nodeLayout(s, counts, flags, leaf_to_internal, num_leaves, first, last, layout, workspace,
                      workspace_bytes);
}

size_t hipcstone_extract_marked_workspace_size(int32_t num_elements) { return extractWorkspace(num_elements); }

#define HIPCSTONE_DEF_EXTRACT(VS, V)                                                                                   \
    hipcstone_status hipcstone_extract_marked_elements_##VS(hipcstone_stream s, const V* source,                      \
                                                            const uint32_t* flags, int32_t first, int32_t second,     \
                                                            V* pairs, uint32_t max_pairs, uint32_t* num_pairs,        \
                                                            void* workspace, size_t workspace_bytes)                  \
    {                                                                                                                  \
        return extractMarked<V>(s, source, flags, first, second, pairs, max_pairs, num_pairs, workspace,              \
                                workspace_bytes);                                                                      \
    }
HIPCSTONE_DEF_EXTRACT(u32, uint32_t)
HIPCSTONE_DEF_EXTRACT(u64, uint64_t)

hipcstone_status hipcstone_gather_ranges(hipcstone_stream s, const uint32_t* range_scan,
                                         const uint32_t* range_offsets, int32_t num_ranges, const void* src,
                                         void* buffer, size_t buffer_size, size_t elem_bytes)
{
    return gatherRanges(s, range_scan, range_offsets, num_ranges, src, buffer, buffer_size, elem_bytes);
}

#define HIPCSTONE_DEF_GATHER_LEAVES(TS, T)                                                                             \
    hipcstone_status hipcstone_gather_leaf_ranges_##TS(hipcstone_stream s, const uint32_t* src_layout,                \
                                                       const uint32_t* dst_layout, int32_t num_leaves, const T* src0, \
                                                       const T* src1, const T* src2, const T* src3, T* dst0, T* dst1, \
                                                       T* dst2, T* dst3)                                               \
    {                                                                                                                  \
        return gatherLeafRanges<T>(s, src_layout, dst_layout, num_leaves, src0, src1, src2, src3, dst0, dst1, dst2,   \
                                   dst3);                                                                              \
    }
HIPCSTONE_DEF_GATHER_LEAVES(f32, float)
HIPCSTONE_DEF_GATHER_LEAVES(f64, double)
HIPCSTONE_DEF_GATHER_LEAVES(u32, uint32_t)
HIPCSTONE_DEF_GATHER_LEAVES(u64, uint64_t)

} // extern "C"
