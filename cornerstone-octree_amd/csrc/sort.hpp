// Internal (C++) entry of the radix sort for the orchestration objects: stable sort of the key bits
// [beginBit, endBit) carrying the permutation iota(0..n), with x,y,z,h gathered by it in the last pass (the fused
// sort + gather of hipcstone_sort_gather_*), and its workspace size.
#pragma once

#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

template<class K>
size_t sortGatherWorkspaceBytes(size_t n);

template<class K, class T>
hipcstone_status sortGatherBits(hipStream_t s, K* keys, uint32_t* perm, size_t n, const T* xIn, const T* yIn,
                                const T* zIn, const T* hIn, T* xOut, T* yOut, T* zOut, T* hOut, K* keyBuf,
                                uint32_t* valueBuf, void* ws, size_t wsBytes, int beginBit, int endBit);

} // namespace hipcstone
