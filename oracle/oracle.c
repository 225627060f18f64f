/*
 * ORACLE — test infrastructure, not product code.
 *
 * A CPU restatement (C11 + OpenMP) of the reference's single-domain octree pipeline
 * (sekelle/cornerstone-octree, include/cstone/): SFC keys, stable sort, cornerstone leaf array
 * (counts / rebalance / update / computeOctree), internal-tree linking, geometric node centers,
 * per-leaf search boxes, halo discovery and radius neighbor search.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and only as
 * the checker. Every function names the reference file:line it restates.
 *
 * Pinning (each function is checked, in tests/test_oracle.py, against the reference's own known-answer tests,
 * against oracle/_ref (the reference headers compiled from /root/reference by oracle/Makefile) when present,
 * and against brute force / closed forms):
 *   keys      morton KATs (test/unit/sfc/morton.cpp:26-73), hilbert octant map, round trip (hilbert.cpp:83-188),
 *             _ref keys on random points (all 8 curve x key x float variants)
 *   sort      stable-sort definition (ties by index), _ref permutation
 *   counts    csarray.cpp:56-80 KAT, brute-force lower_bound differences, _ref
 *   rebalance csarray.cpp:108-122 KAT, _ref computeOctree / updateOctree sequences
 *   link      octree.cpp:99-124 (4x4x4 uniform), connectivity invariants, _ref arrays
 *   centers   closed form from the integer box, _ref
 *   neighbors brute force all-pairs (all_to_all.hpp:26-65) on small inputs, _ref findNeighbors
 *   halos     brute force all-to-all (discovery.cpp:25-50), _ref findHalos
 *   bbox      direct min/max, _ref computeBoundingBox
 *   assign    _ref makeSfcAssignment, equal-count bounds
 *   layout    _ref computeNodeLayout (CPU branch), brute-force kept counts
 *   extract   _ref extractMarkedElements, brute-force run enumeration
 *   gather    brute-force element copy and range enumeration (the reference gatherRanges is GPU-only, not built)
 */

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

#define ORC_MORTON 0
#define ORC_HILBERT 1

/* ---- key templates ---- */
#define KEY uint32_t
#define SKEY int32_t
#define KS u32
#define MAXL 10
#define UNUSED 2
#define KBITS 32
#include "oracle_keys.inc"
#undef KEY
#undef SKEY
#undef KS
#undef MAXL
#undef UNUSED
#undef KBITS

#define KEY uint64_t
#define SKEY int64_t
#define KS u64
#define MAXL 21
#define UNUSED 1
#define KBITS 64
#include "oracle_keys.inc"
#undef KEY
#undef SKEY
#undef KS
#undef MAXL
#undef UNUSED
#undef KBITS

/* ---- coordinate templates ---- */
#define REAL float
#define RS f32
#define FLOOR floorf
#define RINT rintf
#define FABS fabsf
#define SQRT sqrtf
#define REAL_MAX FLT_MAX
#include "oracle_real.inc"
#undef REAL
#undef RS
#undef FLOOR
#undef RINT
#undef FABS
#undef SQRT
#undef REAL_MAX

#define REAL double
#define RS f64
#define FLOOR floor
#define RINT rint
#define FABS fabs
#define SQRT sqrt
#define REAL_MAX DBL_MAX
#include "oracle_real.inc"
#undef REAL
#undef RS
#undef FLOOR
#undef RINT
#undef FABS
#undef SQRT
#undef REAL_MAX

/* ---- key x coordinate templates ---- */
#define KEY uint32_t
#define KS u32
#define MAXL 10
#define REAL float
#define RS f32
#define FLOOR floorf
#define RINT rintf
#define FABS fabsf
#define SQRT sqrtf
#include "oracle_keyreal.inc"
#undef REAL
#undef RS
#undef FLOOR
#undef RINT
#undef FABS
#undef SQRT
#define REAL double
#define RS f64
#define FLOOR floor
#define RINT rint
#define FABS fabs
#define SQRT sqrt
#include "oracle_keyreal.inc"
#undef REAL
#undef RS
#undef FLOOR
#undef RINT
#undef FABS
#undef SQRT
#undef KEY
#undef KS
#undef MAXL

#define KEY uint64_t
#define KS u64
#define MAXL 21
#define REAL float
#define RS f32
#define FLOOR floorf
#define RINT rintf
#define FABS fabsf
#define SQRT sqrtf
#include "oracle_keyreal.inc"
#undef REAL
#undef RS
#undef FLOOR
#undef RINT
#undef FABS
#undef SQRT
#define REAL double
#define RS f64
#define FLOOR floor
#define RINT rint
#define FABS fabs
#define SQRT sqrt
#include "oracle_keyreal.inc"
#undef REAL
#undef RS
#undef FLOOR
#undef RINT
#undef FABS
#undef SQRT
#undef KEY
#undef KS
#undef MAXL

/* layout[0] = 0, layout[i+1] = inclusive scan of counts (test/unit/neighbors/findneighbors.cpp:67-68) */
void orc_leaf_layout(const uint32_t* counts, int numLeaves, uint32_t* layout)
{
    layout[0] = 0;
    for (int i = 0; i < numLeaves; ++i)
        layout[i + 1] = layout[i] + counts[i];
}

/* permutation gather dst[i] = src[perm[i]] (primitives/gather.hpp:90-98) for 4- and 8-byte elements */
void orc_gather_u32(const uint32_t* perm, size_t n, const uint32_t* src, uint32_t* dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        dst[i] = src[perm[i]];
}

void orc_gather_u64(const uint32_t* perm, size_t n, const uint64_t* src, uint64_t* dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i)
        dst[i] = src[perm[i]];
}

/*
 * domain/layout.hpp:153-185 computeNodeLayout (CPU branch :174-182): leaf i keeps its count iff
 * first <= i < last or flags[leafToInternal[i]]; layout = exclusive scan of the kept counts, size numLeaves + 1.
 */
void orc_compute_node_layout(const uint32_t* counts, const uint8_t* flags, const int32_t* leafToInternal,
                             int numLeaves, int first, int last, uint32_t* layout)
{
    uint32_t acc = 0;
    for (int i = 0; i < numLeaves; ++i)
    {
        int have  = (first <= i && i < last) || flags[leafToInternal[i]];
        layout[i] = acc;
        acc += have ? counts[i] : 0u;
    }
    layout[numLeaves] = acc;
}

/*
 * halos/gather_halos_gpu.cu:28-44 gatherRangesKernel, as a loop over ranges: range r covers buffer positions
 * [rangeScan[r], rangeScan[r+1]) (the last one up to bufferSize) and copies them from src starting at
 * rangeOffsets[r]. Elements are elemBytes-byte records.
 */
void orc_gather_ranges(const uint32_t* rangeScan, const uint32_t* rangeOffsets, int numRanges, const void* src,
                       void* buffer, size_t bufferSize, size_t elemBytes)
{
    const unsigned char* s = (const unsigned char*)src;
    unsigned char* b       = (unsigned char*)buffer;
    for (int r = 0; r < numRanges; ++r)
    {
        size_t lo = rangeScan[r];
        size_t hi = r + 1 < numRanges ? rangeScan[r + 1] : bufferSize;
        if (hi > bufferSize) hi = bufferSize;
        for (size_t t = lo; t < hi; ++t)
            memcpy(b + t * elemBytes, s + ((size_t)rangeOffsets[r] + (t - lo)) * elemBytes, elemBytes);
    }
}

/* gather the particles of every kept leaf from the full arrays (srcLayout) to the local arrays (dstLayout) */
void orc_gather_leaf_ranges(const uint32_t* srcLayout, const uint32_t* dstLayout, int numLeaves, const void* src,
                            void* dst, size_t elemBytes)
{
    const unsigned char* s = (const unsigned char*)src;
    unsigned char* d       = (unsigned char*)dst;
    for (int i = 0; i < numLeaves; ++i)
    {
        size_t n = dstLayout[i + 1] - dstLayout[i];
        memcpy(d + (size_t)dstLayout[i] * elemBytes, s + (size_t)srcLayout[i] * elemBytes, n * elemBytes);
    }
}

void orc_free(void* p) { free(p); }
void* orc_malloc(size_t bytes) { return malloc(bytes > 0 ? bytes : 1); }
