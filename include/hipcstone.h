/*
 * hipcstone — MI355X-native (gfx950) single-domain octree pipeline: SFC keys, stable radix sort with permutation
 * gather, cornerstone leaf array build/update, internal-tree linking, geometric node centers, per-leaf search boxes,
 * halo discovery and radius neighbor search.
 *
 * C ABI. All pointers are device pointers unless stated otherwise; every call is stream-ordered on the given
 * stream (NULL = the default stream) and asynchronous unless it returns a host value. No call allocates device
 * memory unless its name says so: scratch comes in through explicit workspace arguments whose size is returned by
 * the matching *_workspace_size() query. Calls are re-entrant: no call keeps state in device globals, so
 * independent calls may run concurrently on different streams.
 *
 * Naming: hipcstone_<function>_<key>_<real> where <key> is u32 (10 levels, 30-bit keys) or u64 (21 levels, 63-bit
 * keys) and <real> is f32 or f64. Each declaration cites the reference interface it replaces
 * (sekelle/cornerstone-octree, paths relative to include/cstone/).
 */
#ifndef HIPCSTONE_H
#define HIPCSTONE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum
{
    HIPCSTONE_OK            = 0,
    HIPCSTONE_INVALID_ARG   = 1, /* sizes or pointers inconsistent; the reference throws std::runtime_error */
    HIPCSTONE_OUT_OF_MEMORY = 2, /* a hipcstone_*_alloc call failed */
    HIPCSTONE_HIP_ERROR     = 3, /* a HIP runtime call failed; see hipcstone_last_error() */
    HIPCSTONE_CAPACITY      = 4, /* an output buffer is smaller than the result; the required size is returned */
    HIPCSTONE_DEVICE_FAULT  = 5  /* a kernel's guard fired: its output is incomplete (hipcstone_device_faults) */
} hipcstone_status;

/* Guards in the kernels that keep every loop bounded. When one fires the kernel still ends normally, its output is
 * not valid, and it sets its bit in the device's fault word; hipcstone_device_faults reads and clears that word.
 * The reference has no counterpart (its kernels spin or walk unboundedly on such input). */
#define HIPCSTONE_FAULT_SORT_LOOKBACK  1u /* a radix sort tile gave up waiting for its predecessor's prefix */
#define HIPCSTONE_FAULT_WALK_BUDGET    2u /* a neighbor walk used up its step budget (a cycle in child offsets) */
#define HIPCSTONE_FAULT_MALFORMED_TREE 4u /* a tree walk met a node index outside the tree, or a stack overflow */
#define HIPCSTONE_FAULT_CAPACITY       8u /* a batched warm step's rebalanced leaf array would exceed its capacity */

/* curve kinds for key encoding (sfc/sfc.hpp:30-40 MortonKey / HilbertKey) */
enum
{
    HIPCSTONE_MORTON  = 0,
    HIPCSTONE_HILBERT = 1
};

/* boundary types (sfc/box.hpp:80-86 BoundaryType) */
enum
{
    HIPCSTONE_OPEN       = 0,
    HIPCSTONE_PERIODIC   = 1,
    HIPCSTONE_FIXED      = 2,
    HIPCSTONE_CUBIC_OPEN = 3
};

/*
 * Coordinate box, the value layout of sfc/box.hpp:95-175 Box<T>: limits {xmin,xmax,ymin,ymax,zmin,zmax},
 * lengths (max-min) and inverse lengths (T(1)/(max-min)) evaluated in T, and the three boundary types.
 * hipcstone_make_box_* fills it exactly as the Box constructor does (box.hpp:100-122).
 */
typedef struct
{
    float lim[6];
    float len[3];
    float ilen[3];
    uint8_t bnd[3];
} hipcstone_box_f32;

typedef struct
{
    double lim[6];
    double len[3];
    double ilen[3];
    uint8_t bnd[3];
} hipcstone_box_f64;

typedef void* hipcstone_stream; /* a hipStream_t */

/* ---- runtime: errors, memory, streams, events (cuda/errorcheck.cuh, cuda/device_vector.h equivalents) ---- */

/* text of the last error on the calling host thread (cuda/errorcheck.cuh:15-27 prints and exits instead) */
const char* hipcstone_last_error(void);
/* Synchronizes s, then reads and clears the fault word of the current device (HIPCSTONE_FAULT_* bits of every guard
 * that fired since the last call, on any stream; *faults may be NULL). Returns HIPCSTONE_DEVICE_FAULT if any bit was
 * set. The pipeline and domain objects also check the word at their own synchronization points. */
hipcstone_status hipcstone_device_faults(hipcstone_stream s, uint32_t* faults);
/* The device address of the current device's fault word (no reference counterpart): for kernels that callers build
 * from this library's device headers (the fused interaction loop of hipcstone_ijloop.hpp), so that their guards
 * raise the same HIPCSTONE_FAULT_* bits that hipcstone_device_faults reports. */
hipcstone_status hipcstone_fault_word(uint32_t** word);
/* Test hook: the number of polls a radix sort tile makes of a predecessor that has not published its prefix before
 * it gives up (HIPCSTONE_FAULT_SORT_LOOKBACK); 0 restores the default (2^22). Applies to later sort launches. */
hipcstone_status hipcstone_set_sort_lookback_limit(uint32_t polls);
/* Tuning knob (no reference counterpart; the reference fixes its traversal tiling at compile time,
 * traversal/find_neighbors.cuh:29-58): waves per workgroup of every later neighbor search launch, 2, 4 (default) or 8
 * (0 restores the default). Results do not depend on it. Process-wide; the autotuner sets it. */
hipcstone_status hipcstone_set_neighbor_waves(int32_t waves);
int32_t hipcstone_get_neighbor_waves(void);
/* library version and the gfx target it was built for */
const char* hipcstone_build_info(void);

hipcstone_status hipcstone_device_alloc(void** ptr, size_t bytes);
hipcstone_status hipcstone_device_free(void* ptr);
hipcstone_status hipcstone_memcpy_h2d(void* dst, const void* src, size_t bytes, hipcstone_stream s);
hipcstone_status hipcstone_memcpy_d2h(void* dst, const void* src, size_t bytes, hipcstone_stream s);
hipcstone_status hipcstone_memcpy_d2d(void* dst, const void* src, size_t bytes, hipcstone_stream s);
/* fill with a byte value through a kernel (capturable in a hipGraph, unlike hipMemsetAsync) */
hipcstone_status hipcstone_fill_bytes(void* dst, uint8_t value, size_t bytes, hipcstone_stream s);
hipcstone_status hipcstone_stream_create(hipcstone_stream* s);
hipcstone_status hipcstone_stream_destroy(hipcstone_stream s);
hipcstone_status hipcstone_stream_sync(hipcstone_stream s);
hipcstone_status hipcstone_device_sync(void);
hipcstone_status hipcstone_event_create(void** ev);
hipcstone_status hipcstone_event_destroy(void* ev);
hipcstone_status hipcstone_event_record(void* ev, hipcstone_stream s);
hipcstone_status hipcstone_event_elapsed_ms(void* start, void* stop, float* ms);
/* host-side Box constructor, sfc/box.hpp:100-122 */
void hipcstone_make_box_f32(float xmin, float xmax, float ymin, float ymax, float zmin, float zmax, int bx, int by,
                            int bz, hipcstone_box_f32* box);
void hipcstone_make_box_f64(double xmin, double xmax, double ymin, double ymax, double zmin, double zmax, int bx,
                            int by, int bz, hipcstone_box_f64* box);

/* ---- a2/a3: SFC keys. Replaces computeSfcKeys(execution::Gpu, x, y, z, keys, n, box), sfc/sfc_gpu.h:24-26 ----
 * keys[i] = sfc3D(x[i], y[i], z[i], box) (sfc/sfc.hpp:142-179). With honor_remove_key != 0, keys equal to
 * removeKey (2^30 / 2^63, tree/definitions.h:70-75) are left unchanged, as the reference kernel does
 * (sfc/sfc_gpu.cu:24-33); with 0 the key array is write-only. */
#define HIPCSTONE_DECL_KEYS(CURVE, KS, K, RS, T)                                                                       \
    hipcstone_status hipcstone_compute_sfc_keys_##CURVE##_##KS##_##RS(hipcstone_stream s, const T* x, const T* y,     \
                                                                      const T* z, K* keys, size_t n,                   \
                                                                      const hipcstone_box_##RS* box,                   \
                                                                      int honor_remove_key);
HIPCSTONE_DECL_KEYS(hilbert, u32, uint32_t, f32, float)
HIPCSTONE_DECL_KEYS(hilbert, u32, uint32_t, f64, double)
HIPCSTONE_DECL_KEYS(hilbert, u64, uint64_t, f32, float)
HIPCSTONE_DECL_KEYS(hilbert, u64, uint64_t, f64, double)
HIPCSTONE_DECL_KEYS(morton, u32, uint32_t, f32, float)
HIPCSTONE_DECL_KEYS(morton, u32, uint32_t, f64, double)
HIPCSTONE_DECL_KEYS(morton, u64, uint64_t, f32, float)
HIPCSTONE_DECL_KEYS(morton, u64, uint64_t, f64, double)

/* ---- a4: stable key sort with uint32 values. Replaces sortByKeyTempStorage / sortByKey(execution::Gpu, first,
 * last, values, keyBuf, valueBuf, tmp, tmpBytes), primitives/primitives_gpu.h:119-130 (CUB SortPairs,
 * primitives_gpu.cu:322-356). Sorts [keys, keys+n) ascending by bits [0, end_bit) in place, carrying values;
 * ties keep their input order. key_buf/value_buf hold n elements each. end_bit = 0 means all bits of the key
 * type. values == NULL sorts keys alone. */
size_t hipcstone_sort_by_key_workspace_size_u32(size_t n);
size_t hipcstone_sort_by_key_workspace_size_u64(size_t n);
hipcstone_status hipcstone_sort_by_key_u32(hipcstone_stream s, uint32_t* keys, uint32_t* values, size_t n,
                                           uint32_t* key_buf, uint32_t* value_buf, void* workspace,
                                           size_t workspace_bytes, int end_bit);
hipcstone_status hipcstone_sort_by_key_u64(hipcstone_stream s, uint64_t* keys, uint32_t* values, size_t n,
                                           uint64_t* key_buf, uint32_t* value_buf, void* workspace,
                                           size_t workspace_bytes, int end_bit);
/* No counterpart in the reference: hipcstone_sort_by_key_u64 without the workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Its workspace is a static device array of the library (about 22.3 MiB per device, held until the
 * process ends), so it accepts at most 2^24 keys (HIPCSTONE_INVALID_ARG beyond) and allocates nothing. Every call on
 * one device shares that device's array: calls that overlap on one device (other streams, other host threads) race
 * and must be serialized by the caller. key_buf/value_buf, values == NULL and end_bit as for
 * hipcstone_sort_by_key_u64. */
hipcstone_status hipcstone_sort_by_key_u64_noalloc(hipcstone_stream s, uint64_t* keys, uint32_t* values, size_t n,
                                                   uint64_t* key_buf, uint32_t* value_buf, int end_bit);

/* ---- a4+a5 fused: keys sorted with the permutation iota(0..n) and x,y,z,h gathered by it, the sequence of
 * sequence() + sortByKey() + gather() x4 (primitives/primitives_gpu.h:100-155, :48-56; domain/domain.hpp:187-188).
 * keys/perm are sorted in place (perm is write-only: the identity is generated on the fly); x,y,z,h are read from
 * the *_in arrays and written in sorted order to *_out. h may be NULL (then h_out is ignored). The workspace also
 * holds 32 bytes per key: when the input keys are far from sorted (more than n/8 descents), x,y,z,h are packed into
 * one record per particle there before the gather. */
size_t hipcstone_sort_gather_workspace_size_u32(size_t n);
size_t hipcstone_sort_gather_workspace_size_u64(size_t n);
#define HIPCSTONE_DECL_SORTGATHER(KS, K, RS, T)                                                                        \
    hipcstone_status hipcstone_sort_gather_##KS##_##RS(                                                                \
        hipcstone_stream s, K* keys, uint32_t* perm, size_t n, const T* x_in, const T* y_in, const T* z_in,           \
        const T* h_in, T* x_out, T* y_out, T* z_out, T* h_out, K* key_buf, uint32_t* value_buf, void* workspace,       \
        size_t workspace_bytes);
HIPCSTONE_DECL_SORTGATHER(u32, uint32_t, f32, float)
HIPCSTONE_DECL_SORTGATHER(u32, uint32_t, f64, double)
HIPCSTONE_DECL_SORTGATHER(u64, uint64_t, f32, float)
HIPCSTONE_DECL_SORTGATHER(u64, uint64_t, f64, double)

/* ---- a5: permutation gather dst[i] = src[perm[i]]. Replaces gather(execution::Gpu, ordering, n, src, dst),
 * primitives/primitives_gpu.h:48-56 ---- */
hipcstone_status hipcstone_gather_u32(hipcstone_stream s, const uint32_t* perm, size_t n, const uint32_t* src,
                                      uint32_t* dst);
hipcstone_status hipcstone_gather_u64(hipcstone_stream s, const uint32_t* perm, size_t n, const uint64_t* src,
                                      uint64_t* dst);

/* ---- a6: node counts. Replaces computeNodeCountsGpu(exec, tree, counts, numNodes, keys, maxCount,
 * useCountsAsGuess), tree/csarray_gpu.h:41-48. counts[i] = min(#keys in [tree[i], tree[i+1]), max_count)
 * (tree/csarray.hpp:181-235); the guess flag only changes the search, never the result. ---- */
hipcstone_status hipcstone_compute_node_counts_u32(hipcstone_stream s, const uint32_t* tree, uint32_t* counts,
                                                   int32_t num_nodes, const uint32_t* keys, size_t num_keys,
                                                   uint32_t max_count, int use_counts_as_guess);
hipcstone_status hipcstone_compute_node_counts_u64(hipcstone_stream s, const uint64_t* tree, uint32_t* counts,
                                                   int32_t num_nodes, const uint64_t* keys, size_t num_keys,
                                                   uint32_t max_count, int use_counts_as_guess);

/* ---- a7: rebalance decision + scan. Replaces computeNodeOpsGpu(exec, tree, numNodes, counts, bucketSize,
 * nodeOps), tree/csarray_gpu.h:61-67: node_ops[0..num_nodes] receives the exclusive scan of the per-node ops
 * (tree/csarray.hpp:267-329); *new_num_nodes (host) = node_ops[num_nodes]; *converged (host) = all ops were 1.
 * Synchronizes the stream. Workspace: hipcstone_node_ops_workspace_size(num_nodes). ---- */
size_t hipcstone_node_ops_workspace_size(int32_t num_nodes);
hipcstone_status hipcstone_compute_node_ops_u32(hipcstone_stream s, const uint32_t* tree, int32_t num_nodes,
                                                const uint32_t* counts, uint32_t bucket_size, int32_t* node_ops,
                                                void* workspace, size_t workspace_bytes, int32_t* new_num_nodes,
                                                int* converged);
hipcstone_status hipcstone_compute_node_ops_u64(hipcstone_stream s, const uint64_t* tree, int32_t num_nodes,
                                                const uint32_t* counts, uint32_t bucket_size, int32_t* node_ops,
                                                void* workspace, size_t workspace_bytes, int32_t* new_num_nodes,
                                                int* converged);

/* ---- a8: rebalance. Replaces rebalanceTreeGpu(exec, tree, numNodes, newNumNodes, nodeOps, newTree),
 * tree/csarray_gpu.h:69-75: new_tree[0..new_num_nodes] from the scanned ops (tree/csarray.hpp:339-389). ---- */
hipcstone_status hipcstone_rebalance_tree_u32(hipcstone_stream s, const uint32_t* tree, int32_t num_nodes,
                                              int32_t new_num_nodes, const int32_t* node_ops, uint32_t* new_tree);
hipcstone_status hipcstone_rebalance_tree_u64(hipcstone_stream s, const uint64_t* tree, int32_t num_nodes,
                                              int32_t new_num_nodes, const int32_t* node_ops, uint64_t* new_tree);

/* ---- a9: one updateOctree step. Replaces updateOctreeGpu(exec, keys, bucketSize, tree, counts, tmpTree,
 * workArray, maxCount), tree/update_gpu.cuh:43-68 (CPU: tree/csarray.hpp:408-426): decision on (tree, counts),
 * rebalance into new_tree, recount into new_counts. *num_nodes is updated (host); returns HIPCSTONE_CAPACITY with
 * *num_nodes set to the required size when capacity (in nodes) is too small, leaving tree untouched.
 * *converged = the input tree's decision was all-unchanged. Synchronizes the stream once.
 * Workspace: hipcstone_update_octree_workspace_size(num_nodes). ---- */
size_t hipcstone_update_octree_workspace_size(int32_t num_nodes);
hipcstone_status hipcstone_update_octree_u32(hipcstone_stream s, const uint32_t* keys, size_t num_keys,
                                             uint32_t bucket_size, const uint32_t* tree, const uint32_t* counts,
                                             int32_t* num_nodes, uint32_t* new_tree, uint32_t* new_counts,
                                             int32_t capacity, uint32_t max_count, void* workspace,
                                             size_t workspace_bytes, int* converged);
hipcstone_status hipcstone_update_octree_u64(hipcstone_stream s, const uint64_t* keys, size_t num_keys,
                                             uint32_t bucket_size, const uint64_t* tree, const uint32_t* counts,
                                             int32_t* num_nodes, uint64_t* new_tree, uint32_t* new_counts,
                                             int32_t capacity, uint32_t max_count, void* workspace,
                                             size_t workspace_bytes, int* converged);

/* ---- a10: internal tree linking. Replaces buildOctreeGpu(exec, cstoneTree, OctreeView), tree/octree_gpu.h:35-45
 * (CPU buildOctreeCpu, tree/octree.hpp:167-196). With L = num_leaves and I = (L-1)/7, N = L+I:
 * prefixes[N], child_offsets[N], parents[max(1,(N-1)/8)], level_range[MAXL+2] (device), internal_to_leaf[N],
 * leaf_to_internal[N]. level_range_host (optional, host) receives a copy of level_range (synchronizes). ---- */
size_t hipcstone_build_octree_workspace_size_u32(int32_t num_leaves);
size_t hipcstone_build_octree_workspace_size_u64(int32_t num_leaves);
hipcstone_status hipcstone_build_octree_u32(hipcstone_stream s, const uint32_t* leaves, int32_t num_leaves,
                                            uint32_t* prefixes, int32_t* child_offsets, int32_t* parents,
                                            int32_t* level_range, int32_t* internal_to_leaf,
                                            int32_t* leaf_to_internal, void* workspace, size_t workspace_bytes,
                                            int32_t* level_range_host);
hipcstone_status hipcstone_build_octree_u64(hipcstone_stream s, const uint64_t* leaves, int32_t num_leaves,
                                            uint64_t* prefixes, int32_t* child_offsets, int32_t* parents,
                                            int32_t* level_range, int32_t* internal_to_leaf,
                                            int32_t* leaf_to_internal, void* workspace, size_t workspace_bytes,
                                            int32_t* level_range_host);

/* ---- a11: geometric node centers and half sizes (Vec3 arrays, 3 T per node). Replaces
 * computeGeoCentersGpu(exec, prefixes, numNodes, centers, sizes, box), focus/source_center_gpu.h:94-99.
 * As in the reference (source_center.hpp:145-158 via sfcKey -> SfcKind), node keys are decoded as Hilbert
 * keys. ---- */
#define HIPCSTONE_DECL_CENTERS(KS, K, RS, T)                                                                           \
    hipcstone_status hipcstone_compute_geo_centers_##KS##_##RS(hipcstone_stream s, const K* prefixes,                 \
                                                               int32_t num_nodes, T* centers, T* sizes,                \
                                                               const hipcstone_box_##RS* box);
HIPCSTONE_DECL_CENTERS(u32, uint32_t, f32, float)
HIPCSTONE_DECL_CENTERS(u32, uint32_t, f64, double)
HIPCSTONE_DECL_CENTERS(u64, uint64_t, f32, float)
HIPCSTONE_DECL_CENTERS(u64, uint64_t, f64, double)

/* ---- Focus-tree source centers and MAC radii (the expansion centers of Domain::syncGrav and
 * Domain::updateExpansionCenters, domain/domain.hpp:220-287, :366-372). centers holds 4 T per node: mass center
 * x, y, z and total |m| (setMac replaces the last by the squared MAC radius).
 * hipcstone_leaf_mass_centers replaces computeLeafSourceCenterGpu(exec, x, y, z, m, leafToInternal, numLeaves, layout,
 * centers), focus/source_center_gpu.h:66-74, with the CPU path's summation order (source_center.hpp:72-78, 110-125:
 * a leaf's particles [layout[i], layout[i+1]) in index order, weights |m|): centers[leaf_to_internal[i]] for every
 * leaf i; leaf_to_internal is the leaf part of the map (leafToInternal + numInternalNodes).
 * hipcstone_upsweep_centers replaces upsweepCentersGpu(exec, numLevels, levelRange, childOffsets, centers),
 * source_center_gpu.h:86-90 (CombineSourceCenter, source_center.hpp:82-98): levels num_levels ... 0 of the host array
 * level_range_host (num_levels + 2 entries, as hipcstone_build_octree_* returns it), each internal node the mass
 * center of its 8 children in child order.
 * hipcstone_set_mac replaces setMacGpu(exec, prefixes, numNodes, macSpheres, invTheta, box), source_center_gpu.h:
 * 110-116 (source_center.hpp:127-140 with macs.hpp computeVecMacR2): centers[i][3] = (2 max(geoSize) inv_theta +
 * |center - geoCenter|)^2 where the mass is nonzero, else 0.
 * hipcstone_geo_mac_spheres replaces geoMacSpheresGpu, source_center_gpu.h:102-108 (computeMinMacR2): centers[i] =
 * {geoCenter, (2 max(geoSize) inv_theta)^2}. Node keys are decoded as Hilbert keys (sfcKey -> SfcKind). ---- */
#define HIPCSTONE_DECL_SOURCE_CENTERS(RS, T)                                                                           \
    hipcstone_status hipcstone_leaf_mass_centers_##RS(hipcstone_stream s, const T* x, const T* y, const T* z,         \
                                                      const T* m, const int32_t* leaf_to_internal,                    \
                                                      int32_t num_leaves, const uint32_t* layout, T* centers);        \
    hipcstone_status hipcstone_upsweep_centers_##RS(hipcstone_stream s, int32_t num_levels,                          \
                                                    const int32_t* level_range_host, const int32_t* child_offsets,   \
                                                    T* centers);
HIPCSTONE_DECL_SOURCE_CENTERS(f32, float)
HIPCSTONE_DECL_SOURCE_CENTERS(f64, double)
#define HIPCSTONE_DECL_MACS(KS, K, RS, T)                                                                              \
    hipcstone_status hipcstone_set_mac_##KS##_##RS(hipcstone_stream s, const K* prefixes, int32_t num_nodes,          \
                                                   T* centers, float inv_theta, const hipcstone_box_##RS* box);       \
    hipcstone_status hipcstone_geo_mac_spheres_##KS##_##RS(hipcstone_stream s, const K* prefixes, int32_t num_nodes,  \
                                                           T* centers, float inv_theta,                               \
                                                           const hipcstone_box_##RS* box);
HIPCSTONE_DECL_MACS(u32, uint32_t, f32, float)
HIPCSTONE_DECL_MACS(u32, uint32_t, f64, double)
HIPCSTONE_DECL_MACS(u64, uint64_t, f32, float)
HIPCSTONE_DECL_MACS(u64, uint64_t, f64, double)

/* ---- a12: leaf layout. layout[0] = 0, layout[i+1] = sum(counts[0..i]) (the inclusive scan of
 * test/unit/neighbors/findneighbors.cpp:67-68 / focus/octree_focus_mpi.hpp:552-554). ---- */
size_t hipcstone_scan_workspace_size(size_t n);
hipcstone_status hipcstone_leaf_layout(hipcstone_stream s, const uint32_t* counts, int32_t num_leaves,
                                       uint32_t* layout, void* workspace, size_t workspace_bytes);

/* ---- a13: per-leaf search boxes. Replaces computeBoundingBoxGpu(exec, x, y, z, h, layout, first, last, scale,
 * centers, sizes), focus/source_center_gpu.h:40-51 (CPU source_center.hpp:28-43). centers (in: seed, e.g. the
 * leaf's geometric center; out: box center) and sizes are leaf-indexed Vec3 arrays. ---- */
#define HIPCSTONE_DECL_BBOX(RS, T)                                                                                     \
    hipcstone_status hipcstone_compute_bounding_box_##RS(hipcstone_stream s, const T* x, const T* y, const T* z,      \
                                                         const T* h, const uint32_t* layout, int32_t first,           \
                                                         int32_t last, T scale, T* centers, T* sizes);
HIPCSTONE_DECL_BBOX(f32, float)
HIPCSTONE_DECL_BBOX(f64, double)

/* ---- a14: halo discovery. Replaces findHalosGpu(exec, prefixes, childOffsets, parents, nodeCenters, nodeSizes,
 * leaves, searchCenters, searchSizes, box, firstNode, lastNode, collisionFlags), traversal/collisions_gpu.h:46-59
 * (CPU collisions.hpp:68-94). flags (node-indexed) are set to 1 for halo nodes and never cleared. num_nodes is the
 * size of the linked tree (bounds every walk). workspace (at least
 * hipcstone_find_halos_workspace_size(num_nodes, last_node - first_node, sizeof(T)) bytes) holds the walk's counters,
 * frontier and the bounds of the search boxes; a NULL workspace is allocated and freed on the stream
 * (hipMallocAsync). ---- */
size_t hipcstone_find_halos_workspace_size(int32_t num_nodes, int32_t num_local_leaves, int32_t real_bytes);
#define HIPCSTONE_DECL_HALOS(KS, K, RS, T)                                                                             \
    hipcstone_status hipcstone_find_halos_##KS##_##RS(                                                                 \
        hipcstone_stream s, const K* prefixes, const int32_t* child_offsets, const int32_t* parents,                  \
        const T* node_centers, const T* node_sizes, const K* leaves, const T* search_centers, const T* search_sizes,   \
        const hipcstone_box_##RS* box, int32_t first_node, int32_t last_node, int32_t num_nodes, uint8_t* flags,       \
        void* workspace, size_t workspace_bytes);
HIPCSTONE_DECL_HALOS(u32, uint32_t, f32, float)
HIPCSTONE_DECL_HALOS(u32, uint32_t, f64, double)
HIPCSTONE_DECL_HALOS(u64, uint64_t, f32, float)
HIPCSTONE_DECL_HALOS(u64, uint64_t, f64, double)

/* ---- SFC assignment. Replaces makeSfcAssignment(numRanks, counts, tree), domain/domaindecomp.hpp:103-118, with
 * uniformBins :32-55: bins of equal particle count over the leaf array, from a uint64 scan of the leaf counts.
 * Outputs are device arrays: tree_offsets[num_ranks+1] (first leaf of each rank, SfcAssignment::treeOffsets),
 * boundaries[num_ranks+1] = leaves[tree_offsets] (the rank boundary keys) and rank_counts[num_ranks] (particles per
 * rank, uint32 as in the reference). Workspace: hipcstone_sfc_assignment_workspace_size(num_leaves). ---- */
size_t hipcstone_sfc_assignment_workspace_size(int32_t num_leaves);
#define HIPCSTONE_DECL_ASSIGN(KS, K)                                                                                   \
    hipcstone_status hipcstone_make_sfc_assignment_##KS(hipcstone_stream s, const K* leaves, const uint32_t* counts,  \
                                                        int32_t num_leaves, int32_t num_ranks, int32_t* tree_offsets, \
                                                        K* boundaries, uint32_t* rank_counts, void* workspace,         \
                                                        size_t workspace_bytes);
HIPCSTONE_DECL_ASSIGN(u32, uint32_t)
HIPCSTONE_DECL_ASSIGN(u64, uint64_t)

/* ---- Halo discovery of one rank. Replaces FocusedOctree::discoverHalos(x, y, z, h, layout, searchExtFact, scratch,
 * accumulate), focus/octree_focus_mpi.hpp:511-568: search boxes of the leaves [first_node, last_node) are seeded with
 * the leaves' geometric centers (node_centers gathered by leaf_to_internal, the leaf part of the map:
 * leafToInternal + numInternalNodes, tree/octree.hpp:252), grown to the AABB of the spheres of
 * radius 2*search_ext*h of their particles (computeBoundingBox), flags are zeroed unless accumulate, and findHalos
 * marks every node that overlaps a search box and lies outside [leaves[first_node], leaves[last_node]).
 * Difference: layout holds the absolute offsets of every leaf's particles in x..h (the reference scans the counts of
 * the assigned leaves into a layout relative to the first assigned particle). search_centers / search_sizes are
 * leaf-indexed Vec3 arrays; only [first_node, last_node) is written. workspace / workspace_bytes as for
 * hipcstone_find_halos_*. ---- */
#define HIPCSTONE_DECL_DISCOVER(KS, K, RS, T)                                                                          \
    hipcstone_status hipcstone_discover_halos_##KS##_##RS(                                                             \
        hipcstone_stream s, const K* prefixes, const int32_t* child_offsets, const int32_t* parents,                  \
        const int32_t* leaf_to_internal, const T* node_centers, const T* node_sizes, const K* leaves,                  \
        int32_t num_nodes, const T* x, const T* y, const T* z, const T* h, const uint32_t* layout,                     \
        int32_t first_node, int32_t last_node, const hipcstone_box_##RS* box, float search_ext, T* search_centers,     \
        T* search_sizes, uint8_t* flags, int32_t accumulate, void* workspace, size_t workspace_bytes);
HIPCSTONE_DECL_DISCOVER(u32, uint32_t, f32, float)
HIPCSTONE_DECL_DISCOVER(u32, uint32_t, f64, double)
HIPCSTONE_DECL_DISCOVER(u64, uint64_t, f32, float)
HIPCSTONE_DECL_DISCOVER(u64, uint64_t, f64, double)

/* ---- a15: local particle layout. Replaces computeNodeLayout(exec, focusLeafCounts, flags, leafToInternal, idx,
 * layout), domain/layout.hpp:153-185: leaf i keeps counts[i] particles iff first <= i < last or
 * flags[leaf_to_internal[i]] != 0, else 0 (leaf_to_internal is the leaf part of the map,
 * tree/octree.hpp:372-375); layout[num_leaves+1] is the exclusive scan of the kept counts, the total in
 * layout[num_leaves]. Workspace: hipcstone_scan_workspace_size(num_leaves). ---- */
hipcstone_status hipcstone_compute_node_layout(hipcstone_stream s, const uint32_t* counts, const uint8_t* flags,
                                               const int32_t* leaf_to_internal, int32_t num_leaves, int32_t first,
                                               int32_t last, uint32_t* layout, void* workspace,
                                               size_t workspace_bytes);

/* ---- a15: marked ranges. Replaces extractMarkedElements(source, flags, firstReqIdx, secondReqIdx),
 * domain/layout.hpp:111-141 (a host std::vector there): every maximal run [a, b) of indices in [first, second) with
 * flags[i+1] > flags[i] gives the pair (source[a], source[b]), in order. flags must be nondecreasing (a layout).
 * pairs receives min(num_pairs, max_pairs) pairs; *num_pairs (device) receives the number of runs. ---- */
size_t hipcstone_extract_marked_workspace_size(int32_t num_elements);
#define HIPCSTONE_DECL_EXTRACT(VS, V)                                                                                  \
    hipcstone_status hipcstone_extract_marked_elements_##VS(hipcstone_stream s, const V* source,                      \
                                                            const uint32_t* flags, int32_t first, int32_t second,     \
                                                            V* pairs, uint32_t max_pairs, uint32_t* num_pairs,        \
                                                            void* workspace, size_t workspace_bytes);
HIPCSTONE_DECL_EXTRACT(u32, uint32_t)
HIPCSTONE_DECL_EXTRACT(u64, uint64_t)

/* ---- a15: range gather. Replaces gatherRanges(exec, rangeScan, rangeOffsets, numRanges, src, buffer, bufferSize),
 * halos/gather_halos_gpu.h:24-31 (kernel gather_halos_gpu.cu:28-44): buffer[t] = src[range_offsets[r] + t -
 * range_scan[r]] with r the last range whose range_scan[r] <= t. Elements are opaque records of elem_bytes in
 * {1, 2, 4, 8, 12, 16} (the reference instantiates int and float arrays of 1..4). range_scan is nondecreasing with
 * range_scan[0] = 0; buffer_size < 2^32. ---- */
hipcstone_status hipcstone_gather_ranges(hipcstone_stream s, const uint32_t* range_scan,
                                         const uint32_t* range_offsets, int32_t num_ranges, const void* src,
                                         void* buffer, size_t buffer_size, size_t elem_bytes);

/* ---- a15: fused halo gather (no reference counterpart as one call; it is extractMarkedElements + gatherRanges
 * over the leaves kept by computeNodeLayout, for up to four arrays at once): for every leaf, the
 * dst_layout[i+1] - dst_layout[i] particles starting at src_layout[i] are copied to dst_layout[i]. src_layout is the
 * leaf layout of the full arrays, dst_layout the one from hipcstone_compute_node_layout. NULL source arrays are
 * skipped. ---- */
#define HIPCSTONE_DECL_GATHER_LEAVES(TS, T)                                                                            \
    hipcstone_status hipcstone_gather_leaf_ranges_##TS(hipcstone_stream s, const uint32_t* src_layout,                \
                                                       const uint32_t* dst_layout, int32_t num_leaves, const T* src0, \
                                                       const T* src1, const T* src2, const T* src3, T* dst0, T* dst1, \
                                                       T* dst2, T* dst3);
HIPCSTONE_DECL_GATHER_LEAVES(f32, float)
HIPCSTONE_DECL_GATHER_LEAVES(f64, double)
HIPCSTONE_DECL_GATHER_LEAVES(u32, uint32_t)
HIPCSTONE_DECL_GATHER_LEAVES(u64, uint64_t)

/* ---- a16: radius neighbor search with the CPU semantics of findNeighbors(x, y, z, h, firstId, lastId, box,
 * OctreeNsView, ngmax, neighbors, neighborsCount), findneighbors.hpp:156-177 — the reference has no GPU free
 * function with this contract; its GPU search is GpuAlwaysTraverseNeighborhoodBuilder
 * (traversal/ijloop/gpu_alwaystraverse.cuh:174-204). For i in [first_id, last_id): row (i-first_id) of neighbors
 * (stride ngmax) receives the ascending indices j != i with |x_j - x_i|^2 < 4 h_i^2 among the leaves the
 * particle's own stackless traversal visits; neighbors_count[i-first_id] is the untruncated count.
 * ext = OctreeNsView::searchExtFactor (tree/octree.hpp:281). neighbors may be NULL when ngmax == 0. ---- */
#define HIPCSTONE_DECL_NEIGHBORS(RS, T)                                                                                \
    hipcstone_status hipcstone_find_neighbors_##RS(                                                                    \
        hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t first_id, uint32_t last_id,      \
        const hipcstone_box_##RS* box, const int32_t* child_offsets, const int32_t* parents,                          \
        const int32_t* internal_to_leaf, const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes,  \
        float ext, uint32_t ngmax, uint32_t* neighbors, uint32_t* neighbors_count);
HIPCSTONE_DECL_NEIGHBORS(f32, float)
HIPCSTONE_DECL_NEIGHBORS(f64, double)

/* ---- the same search with the neighbor rows stored slot-major per group of 64 bodies: for body i, r = i - first_id,
 * neighbor k is at neighbors[((r / 64) * ngmax + k) * 64 + r % 64]; neighbors must hold ceil((last_id - first_id) /
 * 64) * 64 * ngmax elements. No reference counterpart: it is the scratch layout of the ijloop consumer
 * (hipcstone_ijloop.hpp), whose one-thread-per-body fold reads one slot of 64 bodies per coalesced load, like the
 * per-warp column layout of the reference's GpuAlwaysTraverseNeighborhood (gpu_alwaystraverse.cuh:47-75). Counts
 * are exact up to 2^26 neighbors per body. ---- */
#define HIPCSTONE_DECL_NEIGHBORS_SLOT_MAJOR(RS, T)                                                                     \
    hipcstone_status hipcstone_find_neighbors_slot_major_##RS(                                                         \
        hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t first_id, uint32_t last_id,      \
        const hipcstone_box_##RS* box, const int32_t* child_offsets, const int32_t* parents,                          \
        const int32_t* internal_to_leaf, const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes,  \
        float ext, uint32_t ngmax, uint32_t* neighbors, uint32_t* neighbors_count);
HIPCSTONE_DECL_NEIGHBORS_SLOT_MAJOR(f32, float)
HIPCSTONE_DECL_NEIGHBORS_SLOT_MAJOR(f64, double)

/* ---- ijloop (the step after the neighbor search). HIP C++ callers include hipcstone_ijloop.hpp, the counterpart of
 * cstone::ijloop's GpuAlwaysTraverseNeighborhoodBuilder (traversal/ijloop/gpu_alwaystraverse.cuh:174-204, interface
 * ijloop.hpp:29-146), and run their own interaction functors. This C entry runs one fixed interaction through it
 * (ijloop_examples.hip): per body i in [first, last), over i itself and its first ngmax neighbors in ascending order,
 * with w = (1 - d2 / ((4 h_i) h_i))^3: rho = sum m_j w / ((h_i h_i) h_i), count = 1 + neighbors used,
 * dmin / dmax = min / max of d2 over the neighbors, fx = sum of m_j (x_i - x_j) (minimum image for bodies whose 2h box
 * leaves a periodic domain). It equals the reference CPU ijloop (cpu_alwaystraverse.hpp:90-113) running the same
 * interaction. total_bodies is the length of x..m (>= last); with num_groups > 0 only the bodies of the device ranges
 * [group_start[g], group_end[g]) (ascending, disjoint, inside [first, last)) are computed, through the
 * neighborhood's subgroup(), like the reference's GroupView subgroups. ---- */
#define HIPCSTONE_DECL_IJLOOP_EXAMPLE(RS, T)                                                                           \
    hipcstone_status hipcstone_ijloop_example_##RS(                                                                    \
        hipcstone_stream s, const int32_t* child_offsets, const int32_t* parents, const int32_t* internal_to_leaf,    \
        const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes, const hipcstone_box_##RS* box,   \
        uint32_t first, uint32_t last, const T* x, const T* y, const T* z, const T* h, const T* m, uint32_t ngmax,    \
        float ext, uint32_t total_bodies, uint32_t num_groups, const uint32_t* group_start,                          \
        const uint32_t* group_end, T* rho, uint32_t* count, T* dmin, T* dmax, T* fx);
HIPCSTONE_DECL_IJLOOP_EXAMPLE(f32, float)
HIPCSTONE_DECL_IJLOOP_EXAMPLE(f64, double)

/* ---- plain pipeline: one sync step of the single-domain tree, host orchestration of the entries above.
 * Replaces the sequence Domain<KeyType,T,Gpu>::sync at numRanks=1 performs on its global tree
 * (domain/domain.hpp:169-218 -> domain/assignment.hpp:92-144) followed by octreeProperties() + findNeighbors
 * (domain.hpp:377-392): keys, stable sort + gather of x,y,z,h, computeOctree (first call or full_build) or one
 * warm-start updateOctree step, link, geometric centers, leaf layout, and findNeighbors over all particles when
 * neighbors_count != NULL. Inputs are device arrays of n particles; the sorted copies live in the pipeline (see
 * hipcstone_pipeline_view) and may be passed back as the next step's input. Buffers grow on demand; a step
 * synchronizes the stream once per updateOctree call and once to read the level range. ---- */
typedef struct hipcstone_pipeline hipcstone_pipeline;

typedef struct
{
    size_t num_particles;
    int32_t num_leaves;
    int32_t num_nodes;
    int32_t num_internal;
    int32_t converged_calls; /* tree levels the last full build split (it builds the converged tree directly) */
    void* keys;              /* sorted keys (K) */
    void* perm;              /* sorted-to-input permutation (u32) */
    void* x;                 /* sorted coordinates (T) */
    void* y;
    void* z;
    void* h;
    void* leaves; /* cornerstone leaf keys, num_leaves + 1 (K) */
    void* counts; /* particles per leaf (u32) */
    void* prefixes;
    void* child_offsets;
    void* parents;
    void* internal_to_leaf;
    void* leaf_to_internal;
    void* level_range;
    void* centers; /* geometric node centers, Vec3 per node (T) */
    void* sizes;
    void* layout; /* first particle of each leaf, num_leaves + 1 (u32) */
    int32_t level_range_host[24];
    /* state of the last hipcstone_pipeline_halos call (NULL / 0 before the first) */
    int32_t assigned_first;   /* first leaf of the rank */
    int32_t assigned_last;    /* one past its last leaf */
    size_t num_local;         /* assigned + halo particles */
    void* tree_offsets;       /* first leaf of every rank, num_ranks + 1 (i32) */
    void* halo_flags;         /* node-indexed halo flags (u8) */
    void* search_centers;     /* leaf-indexed search boxes of the assigned leaves, Vec3 (T) */
    void* search_sizes;
    void* local_layout;       /* first local particle of each leaf, num_leaves + 1 (u32) */
} hipcstone_pipeline_view;

hipcstone_status hipcstone_pipeline_create(hipcstone_pipeline** p, int curve, int key_bits, int real_bits,
                                           uint32_t bucket_size, uint32_t ngmax);
hipcstone_status hipcstone_pipeline_destroy(hipcstone_pipeline* p);
/* device memory the pipeline holds (its buffers grow on demand and are kept) */
size_t hipcstone_pipeline_device_bytes(const hipcstone_pipeline* p);
hipcstone_status hipcstone_pipeline_sync(hipcstone_pipeline* p, hipcstone_stream s, const void* x, const void* y,
                                         const void* z, const void* h, size_t n, const void* box, int full_build,
                                         uint32_t* neighbors, uint32_t* neighbors_count);
hipcstone_status hipcstone_pipeline_get_view(const hipcstone_pipeline* p, hipcstone_pipeline_view* v);

/* One rank's share of the synced tree: makeSfcAssignment(num_ranks) over the leaf counts, then for `rank`
 * discoverHalos with search_ext (octree_focus_mpi.hpp:511-568), computeNodeLayout (layout.hpp:153-185) and the halo
 * gather of x,y,z,h into x_out..h_out in SFC order (assigned + halo particles; the packing of gatherRanges,
 * gather_halos_gpu.cu:28-57). *num_local receives the local particle count; with capacity below it (or NULL
 * outputs) nothing is gathered and HIPCSTONE_CAPACITY (resp. OK) is returned. Synchronizes twice (rank range,
 * local count). ---- */
hipcstone_status hipcstone_pipeline_halos(hipcstone_pipeline* p, hipcstone_stream s, int32_t num_ranks, int32_t rank,
                                          float search_ext, void* x_out, void* y_out, void* z_out, void* h_out,
                                          size_t capacity, size_t* num_local);

/* Per-stage timing of sync steps with HIP events recorded on the step's stream (no extra synchronization while
 * enabled). Stages: 0 keys, 1 sort+gather, 2 tree update, 3 link, 4 centers, 5 layout, 6 neighbors, and of
 * hipcstone_pipeline_halos: 7 assignment, 8 halo discovery, 9 node layout, 10 halo gather.
 * hipcstone_pipeline_stage_times synchronizes, then returns, per stage, the summed milliseconds and the number of
 * timed launches since timing was enabled (arrays of HIPCSTONE_PIPELINE_STAGES entries). */
#define HIPCSTONE_PIPELINE_STAGES 11
hipcstone_status hipcstone_pipeline_set_timing(hipcstone_pipeline* p, int enable);
hipcstone_status hipcstone_pipeline_stage_times(hipcstone_pipeline* p, double* ms, int32_t* launches);

/* ---- a18: one-rank Domain::sync. Replaces Domain<KeyType, T, Gpu>(rank 0, 1 rank, bucketSize, bucketSizeFocus,
 * theta, box) and its sync(keys, x, y, z, h, {}, scratch), domain/domain.hpp:169-218: the box is fitted to the
 * particles (periodic / fixed dimensions keep the given limits, cubic_open widens to the largest side;
 * sfc/box_mpi.hpp:51-105) and, after the first call, limited to shrink by at most 5 % per side (sfc/box.hpp:398-415);
 * Hilbert keys, stable sort, particles reordered in place into SFC order; the global leaf array follows the update
 * rule of domain/assignment.hpp:115-123 (root start with counts bucketSize-1; repeat while the largest count exceeds
 * bucketSize on the first call or when a step leaves a count of 8*bucketSize or more); the focus tree
 * (bucketSizeFocus) starts from the root with count bucketSizeFocus+1, is updated until converged on the first call
 * and once per later call (octree_focus_mpi.hpp:585-603, domain.hpp:200-203) with every global leaf key mandatory.
 * With one rank there are no halos: start_index = 0, end_index = n, and the particle layout is the scan of the focus
 * leaf counts. theta sets the MAC radius of the expansion centers (1/theta); the MAC flags it also drives in the
 * reference select nothing at one rank (every node lies in the focus). box points to a
 * hipcstone_box_f32/f64 matching real_bits. keys_out (optional, n keys) receives the sorted keys. ---- */
typedef struct hipcstone_domain hipcstone_domain;

typedef struct
{
    size_t num_particles;
    size_t start_index; /* first assigned particle (Domain::startIndex) */
    size_t end_index;   /* one past the last (Domain::endIndex) */
    int32_t num_global_leaves;
    int32_t num_leaves; /* focus tree */
    int32_t num_nodes;
    int32_t num_internal;
    int32_t global_steps;    /* global update steps of the last sync */
    int32_t focus_steps;     /* focus updates of the last sync */
    int32_t focus_converged; /* the last focus update left the tree unchanged */
    double box_lo[3];        /* box of the last sync */
    double box_hi[3];
    void* keys;          /* sorted keys (K) */
    void* global_leaves; /* global leaf array, num_global_leaves + 1 (K) */
    void* global_counts; /* u32 */
    void* leaves;        /* focus leaves, num_leaves + 1 (K) */
    void* counts;        /* focus leaf counts (u32) */
    void* node_counts;   /* focus node counts, num_nodes (u32) */
    void* prefixes;      /* linked focus tree, as hipcstone_build_octree_* returns it */
    void* child_offsets;
    void* parents;
    void* internal_to_leaf;
    void* leaf_to_internal;
    void* level_range;
    void* centers; /* geometric node centers / sizes, Vec3 per node (T) */
    void* sizes;
    void* layout; /* first particle of each focus leaf, num_leaves + 1 (u32) */
    /* focus-tree expansion centers, 4 T per node (mass center x, y, z, squared vector-MAC radius of 1/theta; 0 for a
     * massless node), after hipcstone_domain_sync_grav / _update_expansion_centers on the current tree, else NULL */
    void* expansion_centers;
} hipcstone_domain_view;

hipcstone_status hipcstone_domain_create(hipcstone_domain** d, int key_bits, int real_bits, uint32_t bucket_size,
                                         uint32_t bucket_size_focus, float theta, const void* box);
hipcstone_status hipcstone_domain_destroy(hipcstone_domain* d);
hipcstone_status hipcstone_domain_sync(hipcstone_domain* d, hipcstone_stream s, void* x, void* y, void* z, void* h,
                                       size_t n, void* keys_out);
hipcstone_status hipcstone_domain_get_view(const hipcstone_domain* d, hipcstone_domain_view* v);
/* Particle properties in the order of the last sync. Replaces the particleProperties tuple of Domain::sync
 * (domain.hpp:169-218, reordered with x,y,z,h by updateLayout :479-537) and reapplySync (domain.hpp:295-324): each of
 * the num_arrays device arrays (n elements of elem_bytes[a] in {1,2,4,8,12,16}, in the order the last sync received
 * x, y, z, h) is reordered in place into that sync's SFC order. ---- */
hipcstone_status hipcstone_domain_gather_properties(hipcstone_domain* d, hipcstone_stream s, void* const* arrays,
                                                    const size_t* elem_bytes, int num_arrays);
/* Domain::syncGrav at one rank. Replaces syncGrav(keys, x, y, z, h, m, {}, scratch), domain/domain.hpp:220-287: the
 * particle order and both trees are those of hipcstone_domain_sync (its MAC-driven refinement selects nothing at one
 * rank), m (n T) is reordered with x, y, z, h, and the focus tree's expansion centers are computed as by
 * hipcstone_domain_update_expansion_centers. ---- */
hipcstone_status hipcstone_domain_sync_grav(hipcstone_domain* d, hipcstone_stream s, void* x, void* y, void* z,
                                            void* h, void* m, size_t n, void* keys_out);
/* Replaces Domain::updateExpansionCenters(x, y, z, m, scratch), domain/domain.hpp:366-372 (FocusedOctree::updateCenters
 * + setMacRadius(1 / theta), focus/octree_focus_mpi.hpp:351-414, :446-453): mass centers of the focus leaves over the
 * domain's layout (the CPU path's summation order), the upsweep, and the squared vector-MAC radius; x, y, z, m are
 * device arrays in the order of the last sync. The result is the view's expansion_centers. ---- */
hipcstone_status hipcstone_domain_update_expansion_centers(hipcstone_domain* d, hipcstone_stream s, const void* x,
                                                           const void* y, const void* z, const void* m);
/* per-stage timing as for the pipeline; stages: 0 box, 1 keys, 2 sort+gather, 3 global tree, 4 focus update,
 * 5 focus counts, 6 centers + layout + copy-out, 7 expansion centers */
#define HIPCSTONE_DOMAIN_STAGES 8
hipcstone_status hipcstone_domain_set_timing(hipcstone_domain* d, int enable);
hipcstone_status hipcstone_domain_stage_times(hipcstone_domain* d, double* ms, int32_t* launches);

/* ---- Batched clouds (no single reference counterpart: the reference runs one cloud per Domain or per pipeline
 * call, with host synchronizations inside every tree update, tree/csarray_gpu.cu:97-100,200-202,226-228). Up to 512
 * independent fp32 clouds, stored back to back in cloud order (cloud c holds counts[c] particles, its own box and
 * bucket size), go through keys -> stable sort + gather -> leaf array -> link -> centers -> layout -> findNeighbors
 * with every stage over all clouds in one launch. Each cloud's keys (32-bit Morton or Hilbert, shifted into 64-bit
 * composite keys (c << 54) | (key << 24)), particle order, leaf array and counts (shifted likewise), layout (offset by
 * the cloud's first particle) and neighbor rows and counts (indices offset likewise) are bit-identical to a plain
 * pipeline (hipcstone_pipeline_*) run on that cloud alone with the same box, bucket size and ngmax. The linked
 * nodes are ordered by internal-node slot rather than level by level (children of a node stay consecutive in child
 * order). hipcstone_batched_build runs the first (full) build; hipcstone_batched_step one warm step: an optional
 * drift (particle i of cloud c moves as hipcstone_synth_displace_f32 with seed + c moves the cloud's particle
 * i - first(c); drift_scale 0 skips it), then one updateOctree step of every cloud. A step allocates nothing and never
 * synchronizes the host, so a caller can capture a sequence of steps in one HIP graph; its leaf arrays have a fixed
 * capacity (leaf_capacity_factor * particles / smallest bucket size, at least twice the first build's leaf count),
 * and a step that would exceed it sets HIPCSTONE_FAULT_CAPACITY in the fault word. ---- */
typedef struct hipcstone_batched hipcstone_batched;

typedef struct
{
    uint32_t num_clouds;
    uint32_t num_waves;
    size_t num_particles;
    int32_t num_leaves; /* of the composite leaf array (all clouds, plus one empty leaf per unused cloud slot) */
    int32_t num_nodes;
    int32_t leaf_capacity;
    void* keys; /* composite keys (u64) */
    void* perm; /* u32, the last sort's permutation (global indices) */
    void* x;
    void* y;
    void* z;
    void* h;
    void* leaves; /* num_leaves + 1 composite keys */
    void* counts; /* u32 */
    void* layout; /* u32, num_leaves + 1 */
    void* prefixes;
    void* child_offsets;
    void* internal_to_leaf;
    void* centers;
    void* sizes;
    void* cloud_roots; /* int32 node index of every cloud's root */
    void* neighbors;   /* u32 rows of ngmax per particle (global indices) */
    void* neighbor_counts;
} hipcstone_batched_view;

hipcstone_status hipcstone_batched_create(hipcstone_batched** b, int curve, uint32_t num_clouds,
                                          const uint64_t* counts, const hipcstone_box_f32* boxes,
                                          const uint32_t* bucket_sizes, uint32_t ngmax, double leaf_capacity_factor);
hipcstone_status hipcstone_batched_destroy(hipcstone_batched* b);
/* x, y, z, h of all clouds in cloud order (host or, with device_pointers, device arrays); synchronizes */
hipcstone_status hipcstone_batched_set_particles(hipcstone_batched* b, hipcstone_stream s, const float* x,
                                                 const float* y, const float* z, const float* h, int device_pointers);
hipcstone_status hipcstone_batched_build(hipcstone_batched* b, hipcstone_stream s);
hipcstone_status hipcstone_batched_step(hipcstone_batched* b, hipcstone_stream s, float drift_scale, uint64_t seed,
                                        uint64_t step);
/* synchronizes s (reads the device-side leaf and node counts) */
hipcstone_status hipcstone_batched_get_view(hipcstone_batched* b, hipcstone_stream s, hipcstone_batched_view* v);
size_t hipcstone_batched_device_bytes(const hipcstone_batched* b);
/* Capture `steps` (even) calls of hipcstone_batched_step (step indices first_step .. first_step + steps - 1) on the
 * created stream s into one executable HIP graph, replacing any earlier one; nothing runs until
 * hipcstone_batched_replay launches it on s. A replay runs the same drifts again from wherever the data stand. */
hipcstone_status hipcstone_batched_capture(hipcstone_batched* b, hipcstone_stream s, int32_t steps, float drift_scale,
                                           uint64_t seed, uint64_t first_step);
hipcstone_status hipcstone_batched_replay(hipcstone_batched* b, hipcstone_stream s);

/* ---- Served clouds (no reference counterpart: the reference has no request loop). hipcstone_serve_run runs a
 * list of requests through a loop with continuous batching: requests are released at their arrival times (realtime)
 * or all at once into one FIFO; a request is admitted when one of `slots` workers is free, its cloud has no running
 * or earlier queued request (each cloud's requests run in order), and, for a build, the device memory it is modelled
 * to need fits hbm_budget next to the resident clouds after evicting least-recently-used idle clouds. Every worker
 * runs its requests on its own HIP stream, so admitted requests overlap on the device. A build (kind
 * HIPCSTONE_SERVE_BUILD) uploads the cloud's n particles (host x, y, z, h; a resident cloud of the same id is
 * replaced) and runs a full pipeline sync (Hilbert or Morton 64-bit keys per `curve`, bucket_size, ngmax, box
 * [box_lo, box_hi]^3); the cloud's pipeline then stays resident. An update (HIPCSTONE_SERVE_UPDATE) displaces the
 * resident cloud's particles by N(0, (sigma h)^2) (hipcstone_synth_displace_normal_f32 with the request's seed and
 * id) and runs one warm step from the cloud's previous tree; an update of an evicted cloud first rebuilds it from the
 * request's x, y, z, h (the cloud's build particles). Results hold arrival, admission, start and end times (seconds
 * from the start of the run), the worker slot, the status, the cloud's leaf count and, with digests on, an
 * order-independent 64-bit digest of the request's keys, neighbor counts and stored neighbor rows. ---- */
enum
{
    HIPCSTONE_SERVE_BUILD  = 0,
    HIPCSTONE_SERVE_UPDATE = 1
};
enum
{
    HIPCSTONE_SERVE_BUILT    = 0,
    HIPCSTONE_SERVE_UPDATED  = 1,
    HIPCSTONE_SERVE_REBUILT  = 2,
    HIPCSTONE_SERVE_REJECTED = 3
};

typedef struct
{
    uint64_t id;
    uint32_t cloud;
    uint32_t kind;
    uint64_t n;
    const float* x;
    const float* y;
    const float* z;
    const float* h;
    double arrival_s;
    uint64_t seed;
} hipcstone_serve_request;

typedef struct
{
    uint64_t id;
    uint32_t status;
    uint32_t slot;
    double arrival_s;
    double admit_s;
    double start_s;
    double end_s;
    uint64_t digest;
    int32_t num_leaves;
    uint32_t pad;
    uint64_t resident_bytes; /* device bytes of all resident clouds after this request */
} hipcstone_serve_result;

typedef struct
{
    uint64_t hbm_budget;
    int32_t slots;
    uint32_t bucket_size;
    uint32_t ngmax;
    int32_t curve;
    float sigma;
    float box_lo;
    float box_hi;
    int32_t realtime;
    int32_t digests;
} hipcstone_serve_config;

typedef struct
{
    double makespan_s;
    uint64_t evictions;
} hipcstone_serve_stats;

hipcstone_status hipcstone_serve_run(const hipcstone_serve_config* cfg, const hipcstone_serve_request* requests,
                                     size_t num_requests, hipcstone_serve_result* results,
                                     hipcstone_serve_stats* stats);

/* ---- synthetic workload helper (not an octree stage): x += scale*h*u, u uniform in [-1,1)^3 from a counter-based
 * hash of (seed, step, particle, dimension), clamped into the box; the moving input of multi-step runs ---- */
hipcstone_status hipcstone_synth_displace_f32(hipcstone_stream s, float* x, float* y, float* z, const float* h,
                                              size_t n, float scale, uint64_t seed, uint64_t step,
                                              const hipcstone_box_f32* box);
/* the same with a normal displacement: x_i += sigma * h_i * g_i, g_i ~ N(0, 1)^3 (Box-Muller from the counter-based
 * uniforms of (seed, step, i)), clamped into the box */
hipcstone_status hipcstone_synth_displace_normal_f32(hipcstone_stream s, float* x, float* y, float* z,
                                                     const float* h, size_t n, float sigma, uint64_t seed,
                                                     uint64_t step, const hipcstone_box_f32* box);
hipcstone_status hipcstone_synth_displace_f64(hipcstone_stream s, double* x, double* y, double* z, const double* h,
                                              size_t n, double scale, uint64_t seed, uint64_t step,
                                              const hipcstone_box_f64* box);

/* ---- synthetic workload helper (not an octree stage): Gaussian-mixture particles in [0,1)^3 — `clusters` (at most
 * 256) isotropic Gaussians (centres uniform, sigma log-uniform in [0.005, 0.05]) plus a uniform `background`
 * fraction, clipped to [0,1) — with h = 0.5 (target_neighbors / (4/3 pi rho))^(1/3) from the mixture's density rho
 * at the particle (background plus all clusters) for n_total particles. Particle i is element offset + i of
 * counter-based splitmix64 streams, so a large set can be made in slices. ---- */
hipcstone_status hipcstone_synth_gaussian_mixture_f32(hipcstone_stream s, float* x, float* y, float* z, float* h,
                                                      size_t n, uint64_t offset, uint64_t seed, uint32_t clusters,
                                                      double background, double target_neighbors, double n_total);
hipcstone_status hipcstone_synth_gaussian_mixture_f64(hipcstone_stream s, double* x, double* y, double* z, double* h,
                                                      size_t n, uint64_t offset, uint64_t seed, uint32_t clusters,
                                                      double background, double target_neighbors, double n_total);

#ifdef __cplusplus
}
#endif

#endif /* HIPCSTONE_H */
