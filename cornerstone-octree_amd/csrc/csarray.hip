// a6-a9: the cornerstone leaf array — node counts, rebalance decision, rebalance, one update step.
//
// Reference semantics: tree/csarray.hpp:181-235 (computeNodeCounts), :250-290 (siblingAndLevel,
// calculateNodeOp), :309-329 (rebalanceDecision), :339-389 (processNode, rebalanceTree), :408-426 (updateOctree);
// GPU interface tree/csarray_gpu.h:41-75 and tree/update_gpu.cuh:43-68.
//
// Unlike the reference GPU path (three host round trips per update and __device__ globals, csarray_gpu.cu:68,130),
// all per-call state lives in the caller's workspace and an update step synchronizes once, to learn the new size.
#include "common.hpp"
#include "runtime.hpp"
#include "scan.hpp"

namespace hipcstone
{

// primitives/stl.hpp:64-86 lower_bound
template<class K>
__device__ inline size_t lowerBound(const K* __restrict__ keys, size_t n, K value)
{
    size_t first = 0;
    size_t count = n;
    while (count > 0)
    {
        size_t step = count / 2;
        size_t it   = first + step;
        if (keys[it] < value)
        {
            first = it + 1;
            count -= step + 1;
        }
        else
            count = step;
    }
    return first;
}

// counts[i] = min(lower_bound(tree[i+1]) - lower_bound(tree[i]), maxCount): exactly what computeNodeCounts
// produces, including zeros outside the populated range (csarray.hpp:203-209).
template<class K>
__global__ __launch_bounds__(256) void nodeCountsKernel(const K* __restrict__ tree, uint32_t* __restrict__ counts,
                                                        int32_t numNodes, const K* __restrict__ keys, size_t numKeys,
                                                        uint32_t maxCount)
{
    __shared__ size_t pos[257];
    int32_t base = int32_t(blockIdx.x) * 256;
    int32_t i    = base + int32_t(threadIdx.x);
    if (i <= numNodes) pos[threadIdx.x] = lowerBound(keys, numKeys, tree[i]);
    if (threadIdx.x == 255 && i + 1 <= numNodes) pos[256] = lowerBound(keys, numKeys, tree[i + 1]);
    __syncthreads();
    if (i < numNodes)
    {
        size_t c  = pos[threadIdx.x + 1] - pos[threadIdx.x];
        counts[i] = c < size_t(maxCount) ? uint32_t(c) : maxCount;
    }
}

// tree/csarray.hpp:250-290: op = 0 merge (siblingIdx > 0, the 8 siblings sum to <= bucket), 4096/512/64/8 split,
// else 1. flag[0] is cleared if any op != 1.
template<class K>
__global__ __launch_bounds__(256) void nodeOpsKernel(const K* __restrict__ tree, int32_t numNodes,
                                                     const uint32_t* __restrict__ counts, uint32_t bucketSize,
                                                     int32_t* __restrict__ nodeOps, int* __restrict__ convergedFlag)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == numNodes) nodeOps[numNodes] = 0;
    if (i >= numNodes) return;

    constexpr unsigned L = maxLevel<K>();
    K thisNode           = tree[i];
    K range              = tree[i + 1] - thisNode;
    unsigned level       = treeLevel(range);
    int siblingIdx       = -1;
    if (level > 0)
    {
        siblingIdx    = int(octalDigit(thisNode, level));
        bool siblings = tree[i - siblingIdx + 8] == tree[i - siblingIdx] + nodeRange<K>(level - 1);
        if (!siblings) siblingIdx = -1;
    }

    int op = 1;
    bool merged = false;
    if (siblingIdx > 0)
    {
        const uint32_t* g = counts + i - siblingIdx;
        uint64_t parentCount = uint64_t(g[0]) + g[1] + g[2] + g[3] + g[4] + g[5] + g[6] + g[7];
        if (parentCount <= uint64_t(bucketSize))
        {
            op     = 0;
            merged = true;
        }
    }
    if (!merged)
    {
        uint32_t c = counts[i];
        if (c > bucketSize * 512 && level + 3 < L) op = 4096;
        else if (c > bucketSize * 64 && level + 2 < L) op = 512;
        else if (c > bucketSize * 8 && level + 1 < L) op = 64;
        else if (c > bucketSize && level < L) op = 8;
    }
    nodeOps[i] = op;
    if (op != 1) *convergedFlag = 0; // benign race: every writer stores 0
}

// tree/csarray.hpp:339-366 processNode, one old node per thread; the last key is copied by thread numNodes.
template<class K>
__global__ __launch_bounds__(256) void rebalanceKernel(const K* __restrict__ tree, int32_t numNodes,
                                                       const int32_t* __restrict__ nodeOps, K* __restrict__ newTree)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == numNodes) newTree[nodeOps[numNodes]] = tree[numNodes];
    if (i >= numNodes) return;

    K thisNode     = tree[i];
    unsigned level = treeLevel(K(tree[i + 1] - thisNode));
    int32_t op     = nodeOps[i + 1] - nodeOps[i];
    int32_t dst    = nodeOps[i];
    if (op == 1) { newTree[dst] = thisNode; }
    else if (op >= 8)
    {
        unsigned levelDiff = op == 8 ? 1u : log8ceil(K(uint32_t(op)));
        K step             = nodeRange<K>(level + levelDiff);
        for (int32_t s = 0; s < op; ++s)
            newTree[dst + s] = thisNode + K(s) * step;
    }
}

__global__ void setIntKernel(int* p, int v) { *p = v; }
__global__ void copyI32Kernel(const int32_t* src, int32_t* dst) { *dst = *src; }

// workspace: [flag int, new node count int (256 B)][scan partials]
static size_t nodeOpsWorkspace(int32_t numNodes) { return 256 + scan::workspaceBytes(size_t(numNodes) + 1); }

template<class K>
static hipcstone_status computeNodeCounts(hipcstone_stream s, const K* tree, uint32_t* counts, int32_t numNodes,
                                          const K* keys, size_t numKeys, uint32_t maxCount)
{
    if (numNodes <= 0) return invalid("computeNodeCounts: numNodes must be positive");
    if (!tree || !counts || (numKeys && !keys)) return invalid("computeNodeCounts: null pointer");
    nodeCountsKernel<K><<<ceilDiv(size_t(numNodes), 256), 256, 0, stream(s)>>>(tree, counts, numNodes, keys, numKeys,
                                                                                maxCount);
    HC_CHECK_LAUNCH("nodeCountsKernel");
    return HIPCSTONE_OK;
}

// ops -> exclusive scan in nodeOps[0..numNodes]; returns new size and convergence on the host (one sync)
template<class K>
static hipcstone_status computeNodeOps(hipcstone_stream s, const K* tree, int32_t numNodes, const uint32_t* counts,
                                       uint32_t bucketSize, int32_t* nodeOps, void* ws, size_t wsBytes,
                                       int32_t* newNumNodes, int* converged)
{
    if (numNodes <= 0) return invalid("computeNodeOps: numNodes must be positive");
    if (!tree || !counts || !nodeOps || !ws) return invalid("computeNodeOps: null pointer");
    if (wsBytes < nodeOpsWorkspace(numNodes)) return invalid("computeNodeOps: workspace too small");
    hipStream_t st = stream(s);
    int* flag      = static_cast<int*>(ws);
    void* scanWs   = static_cast<char*>(ws) + 256;
    setIntKernel<<<1, 1, 0, st>>>(flag, 1);
    HC_CHECK_LAUNCH("setIntKernel");
    nodeOpsKernel<K><<<ceilDiv(size_t(numNodes) + 1, 256), 256, 0, st>>>(tree, numNodes, counts, bucketSize, nodeOps,
                                                                         flag);
    HC_CHECK_LAUNCH("nodeOpsKernel");
    if (auto e = scan::exclusiveScan<int32_t>(st, nodeOps, nodeOps, size_t(numNodes) + 1, 0, nullptr, scanWs,
                                              wsBytes - 256))
        return e;
    // the flag and the new node count side by side, read back with one copy into pinned memory
    copyI32Kernel<<<1, 1, 0, st>>>(nodeOps + numNodes, reinterpret_cast<int32_t*>(flag) + 1);
    HC_CHECK_LAUNCH("copyI32Kernel");
    uint32_t* pin = pinnedStaging();
    if (!pin) return HIPCSTONE_HIP_ERROR;
    HC_CHECK(hipMemcpyAsync(pin, flag, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HC_CHECK(hipStreamSynchronize(st));
    if (newNumNodes) *newNumNodes = int32_t(pin[1]);
    if (converged) *converged = int(pin[0]);
    return HIPCSTONE_OK;
}

template<class K>
static hipcstone_status rebalanceTree(hipcstone_stream s, const K* tree, int32_t numNodes, int32_t newNumNodes,
                                      const int32_t* nodeOps, K* newTree)
{
    if (numNodes <= 0 || newNumNodes <= 0) return invalid("rebalanceTree: sizes must be positive");
    if (!tree || !nodeOps || !newTree) return invalid("rebalanceTree: null pointer");
    rebalanceKernel<K><<<ceilDiv(size_t(numNodes) + 1, 256), 256, 0, stream(s)>>>(tree, numNodes, nodeOps, newTree);
    HC_CHECK_LAUNCH("rebalanceKernel");
    return HIPCSTONE_OK;
}

// Warm-step node counts from the particle positions of the old leaf boundaries: oldPos[i] = lower_bound(keys,
// tree[i]) for the old leaf array (the bucket offsets of a sort through those leaves, i in [0, numNodes]). Every
// boundary of the rebalanced array is an old boundary (kept or merged nodes, the first child of a split), whose
// position is oldPos, or a split child's start inside old node i, whose lower bound lies in [oldPos[i], oldPos[i+1])
// and is searched there only. One thread per old node writes the positions of the boundaries it produced (the same
// mapping as rebalanceKernel); counts are then differences of consecutive positions, exactly computeNodeCounts'.
template<class K>
__global__ __launch_bounds__(256) void warmBoundaryPosKernel(const K* __restrict__ tree, int32_t numNodes,
                                                             const int32_t* __restrict__ nodeOps,
                                                             const K* __restrict__ newTree,
                                                             const K* __restrict__ keys,
                                                             const uint32_t* __restrict__ oldPos,
                                                             uint32_t* __restrict__ pos)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == numNodes) pos[nodeOps[numNodes]] = oldPos[numNodes];
    if (i >= numNodes) return;
    const int32_t op  = nodeOps[i + 1] - nodeOps[i];
    const int32_t dst = nodeOps[i];
    if (op == 0) return;
    const uint32_t a = oldPos[i];
    pos[dst]         = a;
    if (op >= 8)
    {
        const uint32_t b = oldPos[i + 1];
        for (int32_t c = 1; c < op; ++c)
            pos[dst + c] = a + uint32_t(lowerBound(keys + a, size_t(b - a), newTree[dst + c]));
    }
}

__global__ __launch_bounds__(256) void countsFromPosKernel(const uint32_t* __restrict__ pos, int32_t numNodes,
                                                           uint32_t maxCount, uint32_t* __restrict__ counts)
{
    int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    const uint32_t c = pos[i + 1] - pos[i];
    counts[i]        = c < maxCount ? c : maxCount;
}

// workspace: [nodeOps (numNodes+1) int32][node-ops workspace]
static size_t updateWorkspace(int32_t numNodes)
{
    return ((size_t(numNodes + 1) * sizeof(int32_t) + 255) & ~size_t(255)) + nodeOpsWorkspace(numNodes);
}

template<class K>
static hipcstone_status updateOctree(hipcstone_stream s, const K* keys, size_t numKeys, uint32_t bucketSize,
                                     const K* tree, const uint32_t* counts, int32_t* numNodes, K* newTree,
                                     uint32_t* newCounts, int32_t capacity, uint32_t maxCount, void* ws,
                                     size_t wsBytes, int* converged, const uint32_t* oldPos = nullptr,
                                     uint32_t* posWs = nullptr)
{
    if (!numNodes || *numNodes <= 0) return invalid("updateOctree: numNodes must be positive");
    if (bucketSize == 0) return invalid("updateOctree: bucketSize must be positive");
    int32_t n = *numNodes;
    if (wsBytes < updateWorkspace(n)) return invalid("updateOctree: workspace too small");
    int32_t* nodeOps = static_cast<int32_t*>(ws);
    size_t opsBytes  = (size_t(n + 1) * sizeof(int32_t) + 255) & ~size_t(255);
    void* opsWs      = static_cast<char*>(ws) + opsBytes;
    int32_t newN     = 0;
    int conv         = 0;
    if (auto e = computeNodeOps<K>(s, tree, n, counts, bucketSize, nodeOps, opsWs, wsBytes - opsBytes, &newN, &conv))
        return e;
    if (converged) *converged = conv;
    if (newN > capacity)
    {
        *numNodes = newN;
        hipcstone::setError("updateOctree: capacity too small for the rebalanced tree");
        return HIPCSTONE_CAPACITY;
    }
    if (auto e = rebalanceTree<K>(s, tree, n, newN, nodeOps, newTree)) return e;
    if (oldPos && posWs)
    {
        warmBoundaryPosKernel<K><<<ceilDiv(size_t(n) + 1, 256), 256, 0, stream(s)>>>(tree, n, nodeOps, newTree, keys,
                                                                                      oldPos, posWs);
        HC_CHECK_LAUNCH("warmBoundaryPosKernel");
        countsFromPosKernel<<<ceilDiv(size_t(newN), 256), 256, 0, stream(s)>>>(posWs, newN, maxCount, newCounts);
        HC_CHECK_LAUNCH("countsFromPosKernel");
    }
    else if (auto e = computeNodeCounts<K>(s, newTree, newCounts, newN, keys, numKeys, maxCount)) return e;
    *numNodes = newN;
    return HIPCSTONE_OK;
}

template<class K>
hipcstone_status updateOctreeWarm(hipcstone_stream s, const K* keys, size_t numKeys, uint32_t bucketSize,
                                  const K* tree, const uint32_t* counts, int32_t* numNodes, K* newTree,
                                  uint32_t* newCounts, int32_t capacity, void* ws, size_t wsBytes, int* converged,
                                  const uint32_t* oldPos, uint32_t* posWs)
{
    return updateOctree<K>(s, keys, numKeys, bucketSize, tree, counts, numNodes, newTree, newCounts, capacity,
                           0xFFFFFFFFu, ws, wsBytes, converged, oldPos, posWs);
}

template hipcstone_status updateOctreeWarm<uint32_t>(hipcstone_stream, const uint32_t*, size_t, uint32_t,
                                                     const uint32_t*, const uint32_t*, int32_t*, uint32_t*, uint32_t*,
                                                     int32_t, void*, size_t, int*, const uint32_t*, uint32_t*);
template hipcstone_status updateOctreeWarm<uint64_t>(hipcstone_stream, const uint64_t*, size_t, uint32_t,
                                                     const uint64_t*, const uint32_t*, int32_t*, uint64_t*, uint32_t*,
                                                     int32_t, void*, size_t, int*, const uint32_t*, uint32_t*);

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

size_t hipcstone_node_ops_workspace_size(int32_t num_nodes) { return nodeOpsWorkspace(num_nodes); }
size_t hipcstone_update_octree_workspace_size(int32_t num_nodes) { return updateWorkspace(num_nodes); }

#define HIPCSTONE_DEF_CSARRAY(KS, K)                                                                                   \
    hipcstone_status hipcstone_compute_node_counts_##KS(hipcstone_stream s, const K* tree, uint32_t* counts,          \
                                                        int32_t num_nodes, const K* keys, size_t num_keys,             \
                                                        uint32_t max_count, int use_counts_as_guess)                   \
    {                                                                                                                  \
        (void)use_counts_as_guess; /* counts are exact either way (csarray.hpp:99-133) */                             \
        return computeNodeCounts<K>(s, tree, counts, num_nodes, keys, num_keys, max_count);                           \
    }                                                                                                                  \
    hipcstone_status hipcstone_compute_node_ops_##KS(hipcstone_stream s, const K* tree, int32_t num_nodes,            \
                                                     const uint32_t* counts, uint32_t bucket_size, int32_t* node_ops,  \
                                                     void* workspace, size_t workspace_bytes, int32_t* new_num_nodes,  \
                                                     int* converged)                                                   \
    {                                                                                                                  \
        return computeNodeOps<K>(s, tree, num_nodes, counts, bucket_size, node_ops, workspace, workspace_bytes,       \
                                 new_num_nodes, converged);                                                            \
    }                                                                                                                  \
    hipcstone_status hipcstone_rebalance_tree_##KS(hipcstone_stream s, const K* tree, int32_t num_nodes,              \
                                                   int32_t new_num_nodes, const int32_t* node_ops, K* new_tree)        \
    {                                                                                                                  \
        return rebalanceTree<K>(s, tree, num_nodes, new_num_nodes, node_ops, new_tree);                               \
    }                                                                                                                  \
    hipcstone_status hipcstone_update_octree_##KS(hipcstone_stream s, const K* keys, size_t num_keys,                 \
                                                  uint32_t bucket_size, const K* tree, const uint32_t* counts,         \
                                                  int32_t* num_nodes, K* new_tree, uint32_t* new_counts,               \
                                                  int32_t capacity, uint32_t max_count, void* workspace,               \
                                                  size_t workspace_bytes, int* converged)                              \
    {                                                                                                                  \
        return updateOctree<K>(s, keys, num_keys, bucket_size, tree, counts, num_nodes, new_tree, new_counts,         \
                               capacity, max_count, workspace, workspace_bytes, converged);                            \
    }

HIPCSTONE_DEF_CSARRAY(u32, uint32_t)
HIPCSTONE_DEF_CSARRAY(u64, uint64_t)

} // extern "C"
