// Host-side helpers shared by the C ABI translation units: error capture and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "hipcstone.h"

namespace hipcstone
{

// Record a message for hipcstone_last_error() on this host thread.
void setError(const std::string& msg);

inline hipcstone_status hipFail(hipError_t e, const char* what)
{
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return HIPCSTONE_HIP_ERROR;
}

inline hipcstone_status invalid(const char* what)
{
    setError(std::string("invalid argument: ") + what);
    return HIPCSTONE_INVALID_ARG;
}

inline hipStream_t stream(hipcstone_stream s) { return reinterpret_cast<hipStream_t>(s); }

// The current device's fault word (device address): kernels whose guards fire atomically OR HIPCSTONE_FAULT_* bits
// into it. Resolved once per device; nullptr if the symbol lookup failed (the error text is then set).
uint32_t* faultWord();
// Read back with a stream-ordered copy into host memory (no synchronization); the orchestration objects enqueue it
// before a synchronization they make anyway and then call checkFaults().
hipcstone_status enqueueFaultRead(hipStream_t s, uint32_t* host);
hipcstone_status checkFaults(uint32_t faults);
// A small pinned host buffer of the calling thread (kPinnedWords words, allocated on first use with hipHostMalloc and
// kept until the process ends; nullptr if that failed, the error text is then set). Device-to-host copies into it are
// asynchronous, so several can ride on one stream synchronization. Words [0, 8) are the tree update's, word 8 the
// orchestration objects' fault-word copy.
constexpr int kPinnedWords = 64;
uint32_t* pinnedStaging();
// the radix sort's look-back poll limit (hipcstone_set_sort_lookback_limit)
uint32_t sortLookbackLimit();
// waves per workgroup of the neighbor search (hipcstone_set_neighbor_waves): 2, 4 or 8
int neighborWaves();

// updateOctree (maxCount unlimited) whose node counts come from oldPos[i] = lower_bound(keys, tree[i]), i in
// [0, *numNodes] (the bucket offsets of this step's sort through the same leaves): a bounded search per split child
// instead of one over all keys per node. posWs holds capacity + 1 words.
template<class K>
hipcstone_status updateOctreeWarm(hipcstone_stream s, const K* keys, size_t numKeys, uint32_t bucketSize,
                                  const K* tree, const uint32_t* counts, int32_t* numNodes, K* newTree,
                                  uint32_t* newCounts, int32_t capacity, void* ws, size_t wsBytes, int* converged,
                                  const uint32_t* oldPos, uint32_t* posWs);

// The neighbor search of hipcstone_find_neighbors_* with a workspace for the compact child blocks the fp32 open-box
// walk reads (neighborGroupsBytes(numNodes) bytes; with less, or none, they are allocated on the stream per call).
size_t neighborGroupsBytes(int32_t numNodes);
template<class T>
struct BoxOf;
template<class T>
hipcstone_status findNeighborsWs(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t firstId,
                                 uint32_t lastId, const typename BoxOf<T>::type* box, const int32_t* childOffsets,
                                 const int32_t* parents, const int32_t* internalToLeaf, const uint32_t* layout,
                                 const T* centers, const T* sizes, int32_t numNodes, float ext, uint32_t ngmax,
                                 uint32_t* neighbors, uint32_t* neighborsCount, void* groupsWs, size_t groupsWsBytes);

} // namespace hipcstone

// Return the status of a failed HIP call from the enclosing C ABI function.
#define HC_CHECK(call)                                                                                                 \
    do {                                                                                                               \
        hipError_t e_ = (call);                                                                                        \
        if (e_ != hipSuccess) return ::hipcstone::hipFail(e_, #call);                                                  \
    } while (0)

// Check that the last kernel launch was accepted.
#define HC_CHECK_LAUNCH(name)                                                                                          \
    do {                                                                                                               \
        hipError_t e_ = hipGetLastError();                                                                             \
        if (e_ != hipSuccess) return ::hipcstone::hipFail(e_, name);                                                   \
    } while (0)
