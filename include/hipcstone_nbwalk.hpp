/*
 * hipcstone neighbor walk: the device side of the radius neighbor search, for HIP C++ code built for gfx950.
 *
 * One wave64 walks the octree for 64 consecutive targets and writes their neighbor rows. The library's search entry
 * points (hipcstone_find_neighbors_*, libhipcstone.so) launch it as is; hipcstone_ijloop.hpp instantiates it in the
 * caller's translation unit inside its fused interaction loop, which folds every wave's rows right after its walk.
 *
 * Reference semantics (findneighbors.hpp:77-177, traversal.hpp:26-69, boxoverlap.hpp:217-250, util/array.hpp:237-240):
 * particle i visits node n iff every node on the root->n path passes
 *     norm2(minDistance(p_i, center_n, size_n)) < 4 h_i^2 * ext^2      (right-fold norm2, CPU minDistance order)
 * and at each visited leaf records j != i with  dx*dx + dy*dy + dz*dz < 4 h_i^2  (left fold), ascending j,
 * keeping the first ngmax and counting all.
 *
 * Design — one wave64 owns 64 consecutive targets (compact in space because they are SFC-ordered):
 *  * Walk: depth-first with an explicit per-wave stack in LDS. Descending into a node evaluates its 8 children at
 *    once (their centers/sizes are contiguous, so a few wide scalar loads fetch them) and stores, per child, the
 *    ballot of lanes whose own predicate chain passed: a lane is tested at a child only if it passed the parent.
 *    A frame holds the children's lane masks and the first child's index; stepping to a child loads its child
 *    offset (and, for a leaf, its leaf index) with scalar loads.
 *    The visited leaf set of every lane is therefore exactly the CPU's, not the superset a group-box walk gives.
 *    Children are taken in order 0..7, so leaves come out in the CPU's visiting order (ascending SFC order).
 *  * Leaves and their lane masks go to a per-wave list, drained in visiting order, so every target's neighbors come
 *    out in ascending j.
 *  * Leaf phase in pair slots: lanes are (target slot, leaf particle) pairs, so nearly all 64 lanes do useful pair
 *    tests even when only a few targets visit a leaf (a per-target lane layout leaves two thirds of the lanes idle
 *    on typical leaves). A ballot ranks each hit among its target's hits in ascending j, so every target's hits of a
 *    leaf are stored as one contiguous run of its row.
 *  * Every loop is bounded: the walk by 2 * numNodes + 8 steps and the stack depth by the key's level count. A guard
 *    that fires ends the walk and ORs its HIPCSTONE_FAULT_* bit into the fault word.
 *
 * Everything is evaluated in the reference's floating-point order; compile with -ffp-contract=off and without
 * packed fp32 ops (-Xclang -target-feature -Xclang -packed-fp32-ops).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>

#include "hipcstone.h"

namespace hipcstone
{

// ---- wave64 primitives ----

constexpr int kWave = 64;

__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// this lane's bit of a wave-uniform mask; the mask is used as a lane predicate as is (no per-lane shift and compare)
__device__ inline bool laneIn(uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

// number of set bits of mask below this lane
__device__ inline unsigned laneRank(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ inline int laneId() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline unsigned uniform(unsigned v) { return unsigned(__builtin_amdgcn_readfirstlane(int(v))); }

__device__ inline uint64_t uniform64(uint64_t v)
{
    uint32_t lo = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v))));
    uint32_t hi = uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(v >> 32))));
    return (uint64_t(hi) << 32) | lo;
}

// order LDS writes of some lanes before LDS reads of other lanes of the same wave
__device__ inline void waveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// XCD-aware work order: workgroups are dealt round-robin to the 8 XCDs (each with its own L2); remap so that each
// XCD walks one contiguous range of groups, whose neighbouring work items share tree nodes in its L2.
// numBlocksPadded is a multiple of kXcds.
constexpr uint32_t kXcds = 8;

__device__ inline uint32_t xcdContiguousBlock(uint32_t b, uint32_t numBlocksPadded)
{
    uint32_t perXcd = numBlocksPadded / kXcds;
    return (b % kXcds) * perXcd + b / kXcds;
}

template<class T>
struct BoxOf;
template<>
struct BoxOf<float>
{
    using type = hipcstone_box_f32;
};
template<>
struct BoxOf<double>
{
    using type = hipcstone_box_f64;
};

namespace nbwalk
{

constexpr int kMaxDepth = 22; // 21 levels below the root for 64-bit keys
constexpr int kLeafList = 64;

// per leaf-chunk size nj in [1, 32]: slots = 64 / nj, the multiplier with floor(l / nj) == (l * magic) >> 16 for
// every lane l < 64 (checked exhaustively), so the per-chunk slot layout needs no division, the lane mask of the
// slots' last lanes {nj - 1, 2 nj - 1, ..., slots * nj - 1} and that of all slot lanes [0, slots * nj); one 32-byte
// scalar load per chunk
struct SlotInfo
{
    uint32_t magic;
    uint32_t slots;
    uint64_t last;
    uint64_t inSlots;
    uint64_t pad;
};
constexpr int kMaxChunk = 32;
struct SlotTable
{
    SlotInfo e[kMaxChunk + 1];
};

constexpr SlotTable makeSlotTable()
{
    SlotTable t{};
    for (int nj = 1; nj <= kMaxChunk; ++nj)
    {
        const int slots = kWave / nj;
        t.e[nj].magic   = (65536u + uint32_t(nj) - 1) / uint32_t(nj);
        t.e[nj].slots   = uint32_t(slots);
        t.e[nj].last    = 0;
        for (int k = 1; k <= slots; ++k)
            t.e[nj].last |= uint64_t(1) << (k * nj - 1);
        t.e[nj].inSlots = slots * nj >= kWave ? ~uint64_t(0) : (uint64_t(1) << (slots * nj)) - 1;
    }
    return t;
}

// one copy per translation unit, internal to it, so that device code addresses it relative to the program counter
// (no global offset table entry to load first)
namespace
{
constexpr SlotTable kSlotTable = makeSlotTable();
}

// v + (this lane's bit of the wave-uniform mask) as one add with the mask as carry-in
__device__ inline uint32_t addLaneBit(uint32_t v, uint64_t mask)
{
    uint32_t r;
    uint64_t carryOut;
    asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(r), "=s"(carryOut) : "v"(v), "s"(mask));
    return r;
}

// all ones if bit b of the wave-uniform v is set, else zero (one scalar sign-extending bit-field extract)
__device__ inline uint32_t sbfeBit(uint32_t v, int b) { return uint32_t(__builtin_amdgcn_sbfe(int(v), b, 1)); }

// a1 where the wave-uniform m is set, a0 elsewhere, as one v_bfi_b32 (the compiler would otherwise turn the bit
// select of an all-ones-or-zero mask into a scalar compare, a scalar select of a lane mask and a v_cndmask)
__device__ inline float pickBits(uint32_t m, uint32_t a1, uint32_t a0)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "s"(m), "v"(a1), "v"(a0));
    return __uint_as_float(r);
}

// bit k set iff m[k] != 0, two scalar instructions per mask (compare, then shift the compare bit in)
__device__ inline uint32_t nonzeroBits8(const uint64_t* m)
{
    uint32_t acc = 0;
    asm("s_cmp_lg_u64 %1, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %2, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %3, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %4, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %5, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %6, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %7, 0\n\ts_addc_u32 %0, %0, %0\n\t"
        "s_cmp_lg_u64 %8, 0\n\ts_addc_u32 %0, %0, %0"
        : "+s"(acc)
        : "s"(m[7]), "s"(m[6]), "s"(m[5]), "s"(m[4]), "s"(m[3]), "s"(m[2]), "s"(m[1]), "s"(m[0])
        : "scc");
    return acc;
}

// the structured buffer store of the LLVM AMDGPU back end (address = base + vindex * stride; with stride != 0 a
// vindex at or past the record count is out of range and the store is dropped), which has no clang builtin
__device__ void structBufferStoreB32(uint32_t data, __amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                     int aux) __asm("llvm.amdgcn.struct.ptr.buffer.store.i32");

// the matching structured buffer load of one 32-bit value (an index at or past the record count reads zero)
__device__ float structBufferLoadF32(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                     int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.f32");

// dword 3 of an untyped, unswizzled buffer resource on gfx9-family targets (32-bit data format); with a stride in
// dword 1, accesses are range checked against the record count in dword 2
constexpr int kBufferFlags = 0x00020000;

template<class T>
struct alignas(4 * sizeof(T)) Particle4
{
    T x, y, z, w;
};

template<class T>
struct alignas(4 * sizeof(T)) TargetRec
{
    Particle4<T> p; // x, y, z, 4h^2
    uint32_t wpos;  // element offset (from the wave's first row) of the target's next neighbor slot
    uint32_t limit; // element offset of the end of the target's row
    uint32_t meta;  // periodic-image flags (bits 0-2) | lane of the target (bits 8-13)
    uint32_t pad;
};

// the LDS of kNbWaves walking waves
template<class T, int kNbWaves>
struct Shared
{
    uint64_t mask[kNbWaves][kMaxDepth][8];
    int32_t co[kNbWaves][kMaxDepth];
    uint32_t rem[kNbWaves][kMaxDepth]; // bit k: child k of the frame has lanes and is not taken yet
    int32_t listLeaf[kNbWaves][kLeafList];
    uint64_t listMask[kNbWaves][kLeafList];
    // the current leaf's active targets, compacted in lane order: wave w's records are [w * 64, w * 64 + 64); the
    // 64 records past the last wave's are padding, so slot indexes up to 127 read inside the array (such reads
    // belong to slots without a target and are masked out) and need no clamp
    TargetRec<T> ctgt[kNbWaves * kWave + kWave];
};

// The 8 children of one internal node in a compact form (fp32 trees): in each dimension the children's centers take
// two values and their sizes one (true of geometric centers, which depend only on a node's integer box), so a
// child's distance terms are one of two values per dimension, computed once per lane for all 8 children. Record g
// describes the children [8 g + 1, 8 g + 9), the child block of some internal node in the level-major layout.
//   v0, v1, s: the two center values and the size per dimension (v1 == v0 where only one value occurs)
//   perm:      bit 4 k + 2 - d is set if child k's center in dimension d is v1
//   flags:     bit k set if child k is internal; kGroupValid if every child matched v0/v1 and s bit for bit (the
//              walk falls back to the children's own centers and sizes otherwise)
//   leaf:      the children's leaf indices (internalToLeaf)
struct alignas(128) ChildGroup
{
    float v0[3], v1[3], s[3];
    uint32_t perm;
    uint32_t flags;
    uint32_t pad[5];
    int32_t leaf[8];
};
typedef uint32_t ChildGroupHead __attribute__((ext_vector_type(16))); // dwords 0-15 of a ChildGroup, one load
constexpr uint32_t kGroupValid = 0x80000000u;

// the ChildGroup records of an fp32 tree with numNodes nodes: (numNodes - 1) / 8 of them (none for a root-only tree)
inline size_t childGroupCount(int32_t numNodes) { return numNodes > 8 ? size_t((numNodes - 1) / 8) : 0; }

// builds the ChildGroup records, one thread per block of 8 children [8 g + 1, 8 g + 9); a template so that every
// translation unit that launches it has its own copy
template<int = 0>
__global__ void childGroupsKernel(const float* __restrict__ centers, const float* __restrict__ sizes,
                                  const int32_t* __restrict__ childOffsets, const int32_t* __restrict__ internalToLeaf,
                                  uint32_t numGroups, ChildGroup* __restrict__ groups)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= numGroups) return;
    const size_t co = size_t(g) * 8 + 1;
    ChildGroup r{};
    bool same      = true;
    uint32_t perm  = 0;
    uint32_t inner = 0;
    for (int d = 0; d < 3; ++d)
    {
        const uint32_t a = __float_as_uint(centers[co * 3 + d]);
        const uint32_t s = __float_as_uint(sizes[co * 3 + d]);
        uint32_t b       = a;
        for (int k = 1; k < 8; ++k)
        {
            const uint32_t c = __float_as_uint(centers[(co + k) * 3 + d]);
            if (c != a)
            {
                if (b == a) b = c;
                same = same && c == b;
                perm |= 1u << (4 * k + 2 - d);
            }
            same = same && __float_as_uint(sizes[(co + k) * 3 + d]) == s;
        }
        r.v0[d] = __uint_as_float(a);
        r.v1[d] = __uint_as_float(b);
        r.s[d]  = __uint_as_float(s);
    }
    for (int k = 0; k < 8; ++k)
    {
        inner |= uint32_t(childOffsets[co + k] != 0) << k;
        r.leaf[k] = internalToLeaf[co + k];
    }
    r.perm    = perm;
    r.flags   = inner | (same ? kGroupValid : 0u);
    groups[g] = r;
}

// a linked octree with its geometry and leaf layout (device pointers; centers and sizes as x,y,z triples). A kernel
// that walks should take these arrays (and x, y, z, h) as separate __restrict__ arguments and assemble the Tree
// itself: only then does the compiler know they are read-only and unaliased and fetch the wave-uniform node values
// with scalar loads (a Tree passed in as a kernel argument costs some 25 vector registers more)
template<class T>
struct Tree
{
    const int32_t* childOffsets;
    const int32_t* internalToLeaf;
    const uint32_t* layout;
    const T* centers;
    const T* sizes;
    int32_t numNodes;
    float ext;
    // optional (fp32, open boxes): the child blocks in compact form, numbered by (first child - 1) / 8
    const ChildGroup* groups = nullptr;
};

/*
 * Walks the tree from `root` for the wave's targets first + lane (lanes with `valid` false take no part) and writes
 * their neighbor rows; returns this lane's neighbor count (all neighbors, of which the first ngmax are stored).
 *
 * Rows: lane l's row starts at rows + l * ngmax (row-major) or, with SlotMajor, neighbor k of lane l is at
 * rows[k * 64 + l]. rowsBytes is the byte size of the row buffer: stores past it are dropped by the hardware's
 * range check (row-major: the rows of the wave's targets only, SlotMajor: all 64 rows).
 * Segmented: the tree is one of many in the same arrays, children may lie before their parents (only the bounds
 * and depth guards and the step budget apply).
 * w is the wave's index in its workgroup (its part of `sh`). Call from all 64 lanes, with wave-uniform arguments
 * except `valid`.
 */
template<class T, bool AnyPbc, bool SlotMajor, bool Segmented, int kNbWaves>
__device__ __attribute__((always_inline)) inline uint32_t
walk(const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, const T* __restrict__ h,
     const typename BoxOf<T>::type& box, const Tree<T>& tree, int32_t root, uint64_t first, bool valid,
     uint32_t ngmax, uint32_t* rowsPtr, uint32_t rowsBytes, Shared<T, kNbWaves>& sh, int w, uint32_t* faults)
{
    const int32_t* __restrict__ childOffsets   = tree.childOffsets;
    const int32_t* __restrict__ internalToLeaf = tree.internalToLeaf;
    const uint32_t* __restrict__ layout        = tree.layout;
    const T* __restrict__ centers              = tree.centers;
    const T* __restrict__ sizes                = tree.sizes;
    const int32_t numNodes                     = tree.numNodes;
    constexpr bool kGroups = std::is_same_v<T, float> && !AnyPbc;
    const ChildGroup* __restrict__ groups = kGroups ? tree.groups : nullptr;

    const int lane   = laneId();
    const uint32_t i = uint32_t(first) + uint32_t(lane);

    T xi = 0, yi = 0, zi = 0, hi = 0;
    if (valid)
    {
        xi = x[i];
        yi = y[i];
        zi = z[i];
        hi = h[i];
    }
    const T radiusSq = T(4.0) * hi * hi;
    const T cellR2   = radiusSq * T(tree.ext) * T(tree.ext);

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

    // minDistance + right-fold norm2 against a node given by its center and half size
    auto nodeDist2 = [&](T c0, T c1, T c2, T s0, T s1, T s2) -> T {
        // the reference's d += |d|; d *= 0.5 is exactly max(d, +0) for finite d (2d and the halving are exact,
        // and d + |d| is +0 for d <= 0)
        T d0 = fmax(fabs(fold(c0 - xi, px, 0)) - s0, T(0));
        T d1 = fmax(fabs(fold(c1 - yi, py, 1)) - s1, T(0));
        T d2 = fmax(fabs(fold(c2 - zi, pz, 2)) - s2, T(0));
        return d0 * d0 + (d1 * d1 + d2 * d2);
    };

    // per-target state: the lane of target t keeps its next row slot in a register (the count is wpos - rowOff, all
    // neighbors counted, only the first ngmax stored); while a leaf is drained the leaf's active targets are copied,
    // compacted, into LDS for the lanes that test them
    constexpr uint32_t kSlotStep = SlotMajor ? uint32_t(kWave) : 1u; // element distance of consecutive slots
    const uint32_t rowOff        = SlotMajor ? uint32_t(lane) : uint32_t(lane) * ngmax;
    const TargetRec<T> myRec{Particle4<T>{xi, yi, zi, radiusSq}, rowOff, rowOff + ngmax * kSlotStep,
                             uint32_t((px ? 1 : 0) | (py ? 2 : 0) | (pz ? 4 : 0)) | (uint32_t(lane) << 8), 0u};
    uint32_t myWpos = rowOff;

    // fp32 particle coordinates as structured buffers of 4-byte records (no range limit): the leaf phase's loads
    // then take the particle index itself, with no 64-bit address arithmetic
    [[maybe_unused]] __amdgpu_buffer_rsrc_t xRes, yRes, zRes;
    if constexpr (std::is_same_v<T, float>)
    {
        xRes = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(x), 4, -1, kBufferFlags);
        yRes = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(y), 4, -1, kBufferFlags);
        zRes = __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(z), 4, -1, kBufferFlags);
    }

    // the wave's output rows as a structured buffer of 4-byte records: a store takes the row slot itself as its
    // index, and a lane without a hit stores at index 2^32 - 1, past the record count, where the hardware's range
    // check drops it (no branch around the store, no byte-offset arithmetic)
    const __amdgpu_buffer_rsrc_t rows = __builtin_amdgcn_make_buffer_rsrc(
        rowsBytes ? static_cast<void*>(rowsPtr) : nullptr, 4, int(rowsBytes / 4u), kBufferFlags);

    // ---- leaf phase. A leaf chunk of nj <= 32 particles is tested against its active targets in pair slots:
    //      lane = (slot, jj) with slot = lane / nj, jj = lane % nj, S = 64 / nj slots; slot s of an iteration tests
    //      active targets k = k0 + s and k0 + S + s (two independent targets per lane, for instruction-level
    //      parallelism) against particle jj. Within a slot, lanes are ascending jj, so the ballot of hits ranks each
    //      hit among its target's hits in ascending j and all of a target's hits of the chunk land contiguously in
    //      its row (coalesced runs). The slot's last lane advances the target's count. Chunks are taken in order,
    //      at most 32 particles each: a leaf of 33-64 particles as one chunk would have one slot per iteration,
    //      as 32 + the rest it has at least two. ----
    int listCount = 0;
    // drain and its pair-loop body are forced inline: as out-of-line calls their captured state (the row
    // buffer resource among it) would live in scratch memory
    auto drain = [&]() __attribute__((always_inline)) {
        waveSync();
        for (int e = 0; e < listCount; ++e)
        {
            int32_t leafIdx = uniform(sh.listLeaf[w][e]);
            uint64_t m      = uniform64(sh.listMask[w][e]);
            uint32_t jb     = uniform(layout[leafIdx]);
            uint32_t je     = uniform(layout[leafIdx + 1]);
            const uint32_t pc     = uint32_t(__builtin_popcountll(m));
            const bool mine       = laneIn(m);
            const unsigned myRank = laneRank(m);
            TargetRec<T>* recs    = &sh.ctgt[w * kWave];
            if (mine)
            {
                TargetRec<T> r = myRec;
                r.wpos         = myWpos;
                recs[myRank]   = r;
            }
            // no active target can reach the end of its row in this leaf (at most je - jb hits each): the
            // per-hit row-end compare is skipped. wpos keeps counting past the row end, so compare in 64 bits.
            const bool fits = uint64_t(myWpos) + uint64_t(je - jb) * kSlotStep <= uint64_t(myRec.limit);
            const bool noOverflow = (ballot(!fits) & m) == 0;
            waveSync();
            for (uint32_t cs = jb, nj = 0; cs < je; cs += nj)
            {
                nj                = je - cs < uint32_t(kMaxChunk) ? je - cs : uint32_t(kMaxChunk);
                const SlotInfo si = kSlotTable.e[nj];
                const uint32_t S  = si.slots;
                uint32_t slot     = __umul24(uint32_t(lane), si.magic) >> 16; // lane < 64, magic <= 2^16
                uint32_t jj       = uint32_t(lane) - slot * nj;
                uint32_t j        = cs + jj;
                // lanes of this and later slots (slot * nj < 64): a hit's rank among its slot's hits
                const uint64_t notBelow = ~uint64_t(0) << (slot * nj);
                const uint32_t nbLo     = uint32_t(notBelow);
                const uint32_t nbHi     = uint32_t(notBelow >> 32);
                // the slots' last lanes: each advances the count of its slot's target
                const uint64_t lastLanes = si.last;
                // lanes of the first b slots, [0, b * nj), as a wave-uniform mask (scalar arithmetic)
                auto slotsBelow = [&](uint32_t b) -> uint64_t {
                    const uint32_t bits = b * nj;
                    return bits >= uint32_t(kWave) ? ~uint64_t(0) : (uint64_t(1) << bits) - 1;
                };
                // lanes past the last slot test a NaN particle, which fails every distance compare
                T xj = std::numeric_limits<T>::quiet_NaN(), yj = xj, zj = xj;
                if constexpr (std::is_same_v<T, float>)
                {
                    // every lane loads (spare lanes a particle of the chunk), only x is replaced by the NaN: no
                    // lane mask around the loads
                    const uint32_t jc = j < je ? j : je - 1;
                    xj                = structBufferLoadF32(xRes, int(jc), 0, 0, 0);
                    yj                = structBufferLoadF32(yRes, int(jc), 0, 0, 0);
                    zj                = structBufferLoadF32(zRes, int(jc), 0, 0, 0);
                    xj                = laneIn(si.inSlots) ? xj : std::numeric_limits<T>::quiet_NaN();
                }
                else
                {
                    if (slot < S)
                    {
                        xj = x[j];
                        yj = y[j];
                        zj = z[j];
                    }
                }
                // j == i is possible only if the chunk overlaps this wave's own targets
                // (one unsigned compare: cs + nj - 1 - first wraps to a large value when the chunk ends before first)
                const bool selfChunk = cs + nj - 1u - uint32_t(first) <= uint32_t(kWave) - 2u + nj;
                const uint32_t jRel  = j - uint32_t(first);

                // the pair test of particle j against a target record: distance (and, in chunks that overlap the
                // wave's own targets, j != target); each condition is a single compare, so its ballot is the
                // compare's own lane mask
                auto close = [&](const TargetRec<T>& rec) -> bool {
                    bool fx = false, fy = false, fz = false;
                    if constexpr (AnyPbc)
                    {
                        fx = rec.meta & 1;
                        fy = rec.meta & 2;
                        fz = rec.meta & 4;
                    }
                    T xx = fold(xj - rec.p.x, fx, 0);
                    T yy = fold(yj - rec.p.y, fy, 1);
                    T zz = fold(zj - rec.p.z, fz, 2);
                    return xx * xx + yy * yy + zz * zz < rec.p.w;
                };
                // slot position of a hit in its target's row: wpos + rank among the slot's hits
                auto position = [&](uint64_t H, uint32_t wpos) -> uint32_t {
                    if constexpr (SlotMajor)
                    {
                        uint32_t r = __builtin_amdgcn_mbcnt_hi(uint32_t(H >> 32) & nbHi,
                                                               __builtin_amdgcn_mbcnt_lo(uint32_t(H) & nbLo, 0u));
                        return wpos + (r << 6);
                    }
                    else
                    {
                        return __builtin_amdgcn_mbcnt_hi(uint32_t(H >> 32) & nbHi,
                                                         __builtin_amdgcn_mbcnt_lo(uint32_t(H) & nbLo, wpos));
                    }
                };
                // the slot position after a lane's own pair: pos, plus one slot if the lane's bit of H is set
                auto advance = [&](uint32_t pos, uint64_t H) -> uint32_t {
                    if constexpr (SlotMajor) { return pos + (laneIn(H) ? kSlotStep : 0u); }
                    else { return addLaneBit(pos, H); }
                };
                auto iterate = [&](auto checkSelf, auto checkLimit) {
                    // hits of particle j against a target record, among the lanes of `in`
                    auto hits = [&](const TargetRec<T>& rec, uint64_t in) -> uint64_t {
                        uint64_t H = ballot(close(rec)) & in;
                        if constexpr (decltype(checkSelf)::value) { H &= ballot(jRel != (rec.meta >> 8)); }
                        return H;
                    };
                    // store index of a lane's pair: its row slot for a hit (inside the row), else 2^32 - 1, past the
                    // record count, where the store is dropped (an inline constant: no register holds it)
                    auto storeOffset = [&](const TargetRec<T>& rec, uint64_t H, uint32_t pos) -> uint32_t {
                        if constexpr (decltype(checkLimit)::value)
                        {
                            return (laneIn(H) & (pos < rec.limit)) ? pos : ~0u;
                        }
                        else { return laneIn(H) ? pos : ~0u; }
                    };
                    // the hits' stores: with the row-end check, every lane stores (a miss or a slot past the row end
                    // at an index past the record count); without it, only the hit lanes store, under their lane
                    // mask (no per-lane index select)
                    auto storeHits = [&](const TargetRec<T>& rec, uint64_t H, uint32_t pos)
                                         __attribute__((always_inline)) {
                        if constexpr (decltype(checkLimit)::value)
                        {
                            structBufferStoreB32(j, rows, int(storeOffset(rec, H, pos)), 0, 0, 0);
                        }
                        else
                        {
                            if (laneIn(H)) structBufferStoreB32(j, rows, int(pos), 0, 0, 0);
                        }
                    };
                    // slot s of an iteration holds targets k0 + s (A) and k0 + S + s (B). In the iterations with
                    // k0 + 2S <= pc every slot holds a target (lanes past the last slot test a NaN particle, which
                    // fails every compare), so no lane mask applies; the last, partial iteration masks the slots
                    // past pc with wave-uniform masks, and has no B half if pc - k0 <= S. Both records are read
                    // before either count is written back.
                    // sameWriteBack: wA == wB (full iterations), both counts are written back under one lane mask
                    auto body2 = [&](TargetRec<T>* rA, uint64_t inA, uint64_t inB, uint64_t wA, uint64_t wB,
                                     auto sameWriteBack) __attribute__((always_inline)) {
                        TargetRec<T>* rB        = rA + S;
                        const TargetRec<T> recA = *rA;
                        const TargetRec<T> recB = *rB;
                        const uint64_t HA       = hits(recA, inA);
                        const uint64_t HB       = hits(recB, inB);
                        const uint32_t posA     = position(HA, recA.wpos);
                        const uint32_t posB     = position(HB, recB.wpos);
                        storeHits(recA, HA, posA);
                        storeHits(recB, HB, posB);
                        if constexpr (decltype(sameWriteBack)::value)
                        {
                            if (laneIn(wA))
                            {
                                rA->wpos = advance(posA, HA);
                                rB->wpos = advance(posB, HB);
                            }
                        }
                        else
                        {
                            if (laneIn(wA)) rA->wpos = advance(posA, HA);
                            if (laneIn(wB)) rB->wpos = advance(posB, HB);
                        }
                    };
                    auto body1 = [&](TargetRec<T>* rA, uint64_t inA, uint64_t wA) __attribute__((always_inline)) {
                        const TargetRec<T> recA = *rA;
                        const uint64_t HA       = hits(recA, inA);
                        const uint32_t posA     = position(HA, recA.wpos);
                        storeHits(recA, HA, posA);
                        if (laneIn(wA)) rA->wpos = advance(posA, HA);
                    };
                    TargetRec<T>* rA = recs + slot; // index < 128: pc <= 64 and slot <= S <= 64
                    uint32_t left    = pc;
                    for (; left >= 2 * S; left -= 2 * S, rA += 2 * S)
                    {
                        body2(rA, ~uint64_t(0), ~uint64_t(0), lastLanes, lastLanes, std::true_type{});
                    }
                    if (left > S)
                    {
                        const uint64_t inB = slotsBelow(left - S);
                        body2(rA, ~uint64_t(0), inB, lastLanes, lastLanes & inB, std::false_type{});
                    }
                    else if (left > 0)
                    {
                        const uint64_t inA = slotsBelow(left);
                        body1(rA, inA, lastLanes & inA);
                    }
                };
                if (selfChunk)
                {
                    if (noOverflow) iterate(std::true_type{}, std::false_type{});
                    else iterate(std::true_type{}, std::true_type{});
                }
                else
                {
                    if (noOverflow) iterate(std::false_type{}, std::false_type{});
                    else iterate(std::false_type{}, std::true_type{});
                }
                waveSync();
            }
            if (mine) myWpos = recs[myRank].wpos;
            waveSync();
        }
        listCount = 0;
    };

    // leaves go to the wave's list (drained before it can overflow: the walk drains it whenever fewer than 9 free
    // entries are left, and one descent appends at most 8 leaves); the drain reads the list after a wave sync
    auto appendLeaf = [&](int32_t leafIdx, uint64_t m) {
        if (lane == 0)
        {
            sh.listLeaf[w][listCount] = leafIdx;
            sh.listMask[w][listCount] = m;
        }
        listCount++;
    };

    // evaluate the 8 children of node range [co, co + 8) for the lanes in `pm`. If every child that some lane
    // passes is a leaf, those leaves are appended in child order and no frame is pushed (true of most nodes at
    // the bottom of the tree); otherwise the children's lane masks are pushed as the frame at `depth`. The 48 node
    // values and the 16 child offsets / leaf indices are loaded up front (contiguous, wave-uniform) and the tests
    // are branch-free. Returns whether a frame was pushed.
    auto pushFrame = [&](int depth, int32_t co, const uint64_t* mk, uint32_t passed) {
        if (lane == 0)
        {
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                sh.mask[w][depth][k] = mk[k];
            }
            sh.co[w][depth]  = co;
            sh.rem[w][depth] = passed;
        }
        waveSync();
    };
    auto descend = [&](int depth, int32_t co, uint64_t pm) -> bool {
        if constexpr (kGroups)
        {
            // compact child block: per lane two distance terms per dimension, each child's terms picked with a
            // bit-field select (v_bfi with a wave-uniform mask); the sums are the per-child nodeDist2 bit for bit
            if (groups != nullptr && ((co - 1) & 7) == 0)
            {
                const ChildGroup* __restrict__ g = groups + ((co - 1) >> 3);
                // geometry, perm and flags arrive with one 64-byte scalar load
                const ChildGroupHead hd = *reinterpret_cast<const ChildGroupHead*>(g);
                const uint32_t flags    = hd[10];
                if (flags & kGroupValid)
                {
                    const uint32_t perm = hd[9];
                    auto term = [&](uint32_t c, uint32_t s, float p) -> uint32_t {
                        float d = fmaxf(fabsf(__uint_as_float(c) - p) - __uint_as_float(s), 0.0f);
                        return __float_as_uint(d * d);
                    };
                    const uint32_t ax0 = term(hd[0], hd[6], xi);
                    const uint32_t ax1 = term(hd[3], hd[6], xi);
                    const uint32_t ay0 = term(hd[1], hd[7], yi);
                    const uint32_t ay1 = term(hd[4], hd[7], yi);
                    const uint32_t az0 = term(hd[2], hd[8], zi);
                    const uint32_t az1 = term(hd[5], hd[8], zi);
                    uint64_t mk[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                    {
                        const float dx = pickBits(sbfeBit(perm, 4 * k + 2), ax1, ax0);
                        const float dy = pickBits(sbfeBit(perm, 4 * k + 1), ay1, ay0);
                        const float dz = pickBits(sbfeBit(perm, 4 * k), az1, az0);
                        mk[k]          = ballot(dx + (dy + dz) < cellR2) & pm;
                    }
                    const uint32_t passed = nonzeroBits8(mk);
                    if ((passed & flags & 0xFFu) == 0)
                    {
#pragma unroll
                        for (int k = 0; k < 8; ++k)
                        {
                            if (passed & (1u << k)) appendLeaf(g->leaf[k], mk[k]);
                        }
                        return false;
                    }
                    pushFrame(depth, co, mk, passed);
                    return true;
                }
            }
        }
        const T* c = centers + size_t(co) * 3;
        const T* s = sizes + size_t(co) * 3;
        T cv[24], sv[24];
        int32_t cc[8], lf[8];
#pragma unroll
        for (int k = 0; k < 24; ++k)
        {
            cv[k] = c[k];
            sv[k] = s[k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            cc[k] = childOffsets[co + k];
            lf[k] = internalToLeaf[co + k];
        }
        uint64_t mk[8];
        // the ballot of the bare compare is the compare's own lane mask; the parent's lanes are applied as a
        // scalar AND
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            T d2  = nodeDist2(cv[3 * k], cv[3 * k + 1], cv[3 * k + 2], sv[3 * k], sv[3 * k + 1], sv[3 * k + 2]);
            mk[k] = ballot(d2 < cellR2) & pm;
        }
        bool allLeaves = true;
#pragma unroll
        for (int k = 0; k < 8; ++k)
        {
            allLeaves = allLeaves && (mk[k] == 0 || cc[k] == 0);
        }
        if (allLeaves)
        {
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                if (mk[k] != 0) appendLeaf(lf[k], mk[k]);
            }
            return false;
        }
        pushFrame(depth, co, mk, nonzeroBits8(mk));
        return true;
    };

    // root (traversal.hpp:35-45)
    const T* rc       = centers + size_t(root) * 3;
    const T* rsz      = sizes + size_t(root) * 3;
    uint64_t rootMask = ballot(valid & (nodeDist2(rc[0], rc[1], rc[2], rsz[0], rsz[1], rsz[2]) < cellR2));
    if (rootMask != 0)
    {
        int32_t rootChild = uniform(childOffsets[root]);
        // a guard that fires (out-of-range node, stack overflow, step budget) ends the walk and sets its bit in
        // the fault word: the wave's rows and counts are then incomplete
        uint32_t fault = 0;
        if (rootChild == 0) { appendLeaf(uniform(internalToLeaf[root]), rootMask); }
        else if (rootChild < 0 || rootChild + 8 > numNodes) { fault = HIPCSTONE_FAULT_MALFORMED_TREE; }
        else
        {
            int depth         = descend(0, rootChild, rootMask) ? 0 : -1;
            // a 32-bit budget (saturated), so that the per-step test is a scalar compare
            const uint64_t budget = 2 * uint64_t(uint32_t(numNodes)) + 8;
            uint32_t stepsLeft    = budget > 0xFFFFFFFFu ? 0xFFFFFFFFu : uint32_t(budget);
            while (depth >= 0)
            {
                if (stepsLeft-- == 0)
                {
                    fault = HIPCSTONE_FAULT_WALK_BUDGET;
                    break;
                }
                if (listCount > kLeafList - 9) drain();
                // the frame's next child with lanes: the lowest bit of its remaining-children word
                const uint32_t rem = uniform(sh.rem[w][depth]);
                if (rem == 0)
                {
                    --depth;
                    continue;
                }
                const int k        = __builtin_ctz(rem);
                const uint64_t m   = uniform64(sh.mask[w][depth][k]);
                int32_t node       = uniform(sh.co[w][depth]) + k;
                if (lane == 0) sh.rem[w][depth] = rem & (rem - 1);
                waveSync();
                if (node <= 0 || node >= numNodes)
                {
                    fault = HIPCSTONE_FAULT_MALFORMED_TREE;
                    break;
                }
                int32_t cco = uniform(childOffsets[node]);
                if (cco == 0) { appendLeaf(uniform(internalToLeaf[node]), m); }
                else
                {
                    if (depth + 1 >= kMaxDepth || (!Segmented && cco <= node) || cco <= 0 || cco + 8 > numNodes)
                    {
                        fault = HIPCSTONE_FAULT_MALFORMED_TREE;
                        break;
                    }
                    if (descend(depth + 1, cco, m)) ++depth;
                }
            }
        }
        if (fault != 0 && lane == 0) atomicOr(faults, fault);
        drain();
    }
    return SlotMajor ? (myWpos - rowOff) >> 6 : myWpos - rowOff;
}

} // namespace nbwalk
} // namespace hipcstone
