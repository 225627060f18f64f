// a16: radius neighbor search with the per-particle semantics of the reference CPU path.
//
// Reference semantics (findneighbors.hpp:77-177, traversal.hpp:26-69, boxoverlap.hpp:217-250, util/array.hpp:237-240):
// particle i visits node n iff every node on the root->n path passes
//     norm2(minDistance(p_i, center_n, size_n)) < 4 h_i^2 * ext^2      (right-fold norm2, CPU minDistance order)
// and at each visited leaf records j != i with  dx*dx + dy*dy + dz*dz < 4 h_i^2  (left fold), ascending j,
// keeping the first ngmax and counting all.
//
// GPU design — one wave64 owns 64 consecutive targets (compact in space because they are SFC-ordered):
//  * Walk: depth-first with an explicit per-wave stack in LDS. Descending into a node evaluates its 8 children at
//    once (their centers/sizes are contiguous, so a few wide scalar loads fetch them) and stores, per child, the
//    ballot of lanes whose own predicate chain passed: a lane is tested at a child only if it passed the parent.
//    A frame holds the children's lane masks and the first child's index; stepping to a child loads its child
//    offset (and, for a leaf, its leaf index) with scalar loads.
//    The visited leaf set of every lane is therefore exactly the CPU's, not the superset a group-box walk gives.
//    Children are taken in order 0..7, so leaves come out in the CPU's visiting order (ascending SFC order).
//  * Leaves and their lane masks go to a per-wave list, drained in visiting order, so every target's neighbors come
//    out in ascending j.
//  * Leaf phase in pair slots: lanes are (target slot, leaf particle) pairs, so nearly all 64 lanes do useful pair
//    tests even when only a few targets visit a leaf (a per-target lane layout leaves two thirds of the lanes idle
//    on typical leaves). A ballot ranks each hit among its target's hits in ascending j, so every target's hits of a
//    leaf are stored as one contiguous run of its row.
//  * Every loop is bounded: the walk by 2 * numNodes + 8 steps and the stack depth by the key's level count.
#include <limits>
#include <type_traits>

#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

// waves per workgroup: 4 by default; 2 and 8 are the other shapes the autotuner measures
// (hipcstone_set_neighbor_waves). Every wave owns its own targets, stack and leaf list, so the shape changes only
// how waves are grouped onto CUs (LDS per workgroup, scheduling granularity), never a result.
constexpr int kMaxDepth  = 22; // 21 levels below the root for 64-bit keys
constexpr int kLeafList  = 64;

// per leaf-chunk size nj in [1, 64]: slots = 64 / nj and the multiplier with floor(l / nj) == (l * magic) >> 16
// for every lane l < 64 (checked exhaustively), so the per-chunk slot layout needs no division
struct SlotTables
{
    uint32_t magic[kWave + 1];
    uint32_t slots[kWave + 1];
};

constexpr SlotTables makeSlotTables()
{
    SlotTables t{};
    for (int nj = 1; nj <= kWave; ++nj)
    {
        t.magic[nj] = (65536u + uint32_t(nj) - 1) / uint32_t(nj);
        t.slots[nj] = uint32_t(kWave / nj);
    }
    return t;
}

__constant__ SlotTables kSlotTables = makeSlotTables();

// v + (this lane's bit of the wave-uniform mask) as one add with the mask as carry-in
__device__ inline uint32_t addLaneBit(uint32_t v, uint64_t mask)
{
    uint32_t r;
    uint64_t carryOut;
    asm("v_addc_co_u32_e64 %0, %1, %2, 0, %3" : "=v"(r), "=s"(carryOut) : "v"(v), "s"(mask));
    return r;
}

// dword 3 of a raw (untyped, unswizzled) buffer resource on gfx9-family targets: 32-bit data format, range checked
// against the byte count in dword 2
constexpr int kRawBufferFlags = 0x00020000;

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

template<class T, int kNbWaves>
struct NbShared
{
    uint64_t mask[kNbWaves][kMaxDepth][8];
    int32_t co[kNbWaves][kMaxDepth];
    int32_t next[kNbWaves][kMaxDepth];
    int32_t listLeaf[kNbWaves][kLeafList];
    uint64_t listMask[kNbWaves][kLeafList];
    // the current leaf's active targets, compacted in lane order: wave w's records are [w * 64, w * 64 + 64); the
    // 64 records past the last wave's are padding, so slot indexes up to 127 read inside the array (such reads
    // belong to slots without a target and are masked out) and need no clamp
    TargetRec<T> ctgt[kNbWaves * kWave + kWave];
};

// SlotMajor: the wave's 64 rows are stored slot-major (neighbor k of the wave's lane l at k * 64 + l) instead of
// row-major (l * ngmax + k), so that a consumer with one thread per target reads each slot with one coalesced load.
// Segmented: many independent trees in one set of arrays (the batched clouds). Wave `group` then takes its targets
// [first, last) (at most 64, inside one cloud) and the index of its cloud's root node from the wave table
// (waves[3 g], waves[3 g + 1], cloud = waves[3 g + 2], root = roots[cloud]); rows and counts are indexed by the
// particle index itself (firstId = 0), and a child block may lie before its parent (the batched link places children
// by internal-node slot, not level by level), so only the bounds and depth guards and the step budget apply.
template<class T, bool AnyPbc, bool SlotMajor, bool Segmented, int kNbWaves>
__global__ __launch_bounds__(kNbWaves * kWave) void findNeighborsKernel(
    const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z, const T* __restrict__ h,
    uint32_t firstId, uint32_t lastId, typename BoxOf<T>::type box, const int32_t* __restrict__ childOffsets,
    const int32_t* __restrict__ internalToLeaf, const uint32_t* __restrict__ layout, const T* __restrict__ centers,
    const T* __restrict__ sizes, int32_t numNodes, float ext, uint32_t ngmax, uint32_t* __restrict__ neighbors,
    uint32_t* __restrict__ neighborsCount, uint32_t numBlocksPadded, uint32_t* __restrict__ faults,
    const uint32_t* __restrict__ waves = nullptr, uint32_t numWaves = 0, const int32_t* __restrict__ roots = nullptr)
{
    __shared__ NbShared<T, kNbWaves> sh;
    const int lane       = laneId();
    const int w          = threadIdx.x / kWave;
    const uint32_t block = xcdContiguousBlock(blockIdx.x, numBlocksPadded);
    const uint32_t group = uniform(block * uint32_t(kNbWaves) + uint32_t(w));
    uint64_t first;
    int32_t root = 0;
    if constexpr (Segmented)
    {
        if (group >= numWaves) return;
        first  = uniform(waves[3 * group]);
        lastId = uniform(waves[3 * group + 1]);
        root   = uniform(roots[uniform(waves[3 * group + 2])]);
        if (root < 0 || root >= numNodes)
        {
            if (lane == 0) atomicOr(faults, HIPCSTONE_FAULT_MALFORMED_TREE);
            return;
        }
    }
    else
    {
        first = uint64_t(firstId) + uint64_t(group) * kWave;
        if (first >= lastId) return;
    }
    const uint32_t i     = uint32_t(first) + uint32_t(lane);
    const bool valid     = i < lastId;
    const size_t rowBase = Segmented ? size_t(first) : size_t(group) * kWave; // row of lane 0 in the output

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

    // the wave's 64 output rows as a raw buffer: stores at an offset past its end are dropped by the hardware's
    // range check, so a lane without a hit stores nowhere instead of branching around the store
    const uint32_t validRows = uint64_t(lastId) - first < uint64_t(kWave) ? uint32_t(uint64_t(lastId) - first) : kWave;
    const uint32_t rowsBytes = ngmax ? (SlotMajor ? uint32_t(kWave) : validRows) * ngmax * uint32_t(sizeof(uint32_t))
                                     : 0u;
    const __amdgpu_buffer_rsrc_t rows =
        __builtin_amdgcn_make_buffer_rsrc(ngmax ? static_cast<void*>(neighbors + rowBase * ngmax) : nullptr, 0,
                                          int(rowsBytes), kRawBufferFlags);

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
            const bool fits =
                uint64_t(myWpos) + uint64_t(je - jb) * kSlotStep <= uint64_t(myRec.limit);
            const bool noOverflow = (ballot(!fits) & m) == 0;
            waveSync();
            for (uint32_t cs = jb, nj = 0; cs < je; cs += nj)
            {
                nj             = je - cs < 32u ? je - cs : 32u;
                uint32_t magic = uniform(kSlotTables.magic[nj]);
                uint32_t S     = uniform(kSlotTables.slots[nj]);
                uint32_t slot  = __umul24(uint32_t(lane), magic) >> 16; // lane < 64, magic <= 2^16
                uint32_t jj    = uint32_t(lane) - slot * nj;
                bool jv        = slot < S;
                uint32_t j     = cs + jj;
                // lanes of this and later slots (slot * nj < 64): a hit's rank among its slot's hits
                const uint64_t notBelow = ~((uint64_t(1) << (slot * nj)) - 1);
                const uint32_t nbLo     = uint32_t(notBelow);
                const uint32_t nbHi     = uint32_t(notBelow >> 32);
                // the slots' last lanes: each advances the count of its slot's target
                const uint64_t lastLanes = ballot(jv && jj == nj - 1);
                // lanes of the first b slots, [0, b * nj), as a wave-uniform mask (scalar arithmetic)
                auto slotsBelow = [&](uint32_t b) -> uint64_t {
                    const uint32_t bits = b * nj;
                    return bits >= uint32_t(kWave) ? ~uint64_t(0) : (uint64_t(1) << bits) - 1;
                };
                // lanes past the last slot test a NaN particle, which fails every distance compare
                T xj = std::numeric_limits<T>::quiet_NaN(), yj = xj, zj = xj;
                if (jv)
                {
                    xj = x[j];
                    yj = y[j];
                    zj = z[j];
                }
                // j == i is possible only if the chunk overlaps this wave's own targets
                const bool selfChunk = cs < uint32_t(first) + uint32_t(kWave) && cs + nj > uint32_t(first);
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
                    // store byte offset of a lane's pair: its row slot for a hit (inside the row), else past the
                    // end of the buffer, where the store is dropped
                    auto storeOffset = [&](const TargetRec<T>& rec, uint64_t H, uint32_t pos) -> uint32_t {
                        if constexpr (decltype(checkLimit)::value)
                        {
                            return (laneIn(H) & (pos < rec.limit)) ? pos * 4u : rowsBytes;
                        }
                        else { return laneIn(H) ? pos * 4u : rowsBytes; }
                    };
                    // slot s of an iteration holds targets k0 + s (A) and k0 + S + s (B). In the iterations with
                    // k0 + 2S <= pc every slot holds a target (lanes past the last slot test a NaN particle, which
                    // fails every compare), so no lane mask applies; the last, partial iteration masks the slots
                    // past pc with wave-uniform masks, and has no B half if pc - k0 <= S. Both records are read
                    // before either count is written back.
                    auto body2 = [&](TargetRec<T>* rA, uint64_t inA, uint64_t inB, uint64_t wA,
                                     uint64_t wB) __attribute__((always_inline)) {
                        TargetRec<T>* rB        = rA + S;
                        const TargetRec<T> recA = *rA;
                        const TargetRec<T> recB = *rB;
                        const uint64_t HA       = hits(recA, inA);
                        const uint64_t HB       = hits(recB, inB);
                        const uint32_t posA     = position(HA, recA.wpos);
                        const uint32_t posB     = position(HB, recB.wpos);
                        __builtin_amdgcn_raw_buffer_store_b32(j, rows, int(storeOffset(recA, HA, posA)), 0, 0);
                        __builtin_amdgcn_raw_buffer_store_b32(j, rows, int(storeOffset(recB, HB, posB)), 0, 0);
                        if (laneIn(wA)) rA->wpos = advance(posA, HA);
                        if (laneIn(wB)) rB->wpos = advance(posB, HB);
                    };
                    auto body1 = [&](TargetRec<T>* rA, uint64_t inA, uint64_t wA) __attribute__((always_inline)) {
                        const TargetRec<T> recA = *rA;
                        const uint64_t HA       = hits(recA, inA);
                        const uint32_t posA     = position(HA, recA.wpos);
                        __builtin_amdgcn_raw_buffer_store_b32(j, rows, int(storeOffset(recA, HA, posA)), 0, 0);
                        if (laneIn(wA)) rA->wpos = advance(posA, HA);
                    };
                    TargetRec<T>* rA = recs + slot; // index < 128: pc <= 64 and slot <= S <= 64
                    uint32_t left    = pc;
                    for (; left >= 2 * S; left -= 2 * S, rA += 2 * S)
                    {
                        body2(rA, ~uint64_t(0), ~uint64_t(0), lastLanes, lastLanes);
                    }
                    if (left > S)
                    {
                        const uint64_t inB = slotsBelow(left - S);
                        body2(rA, ~uint64_t(0), inB, lastLanes, lastLanes & inB);
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
    auto descend = [&](int depth, int32_t co, uint64_t pm) -> bool {
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
        if (lane == 0)
        {
#pragma unroll
            for (int k = 0; k < 8; ++k)
            {
                sh.mask[w][depth][k] = mk[k];
            }
            sh.co[w][depth]   = co;
            sh.next[w][depth] = 0;
        }
        waveSync();
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
            int64_t stepsLeft = 2 * int64_t(numNodes) + 8;
            while (depth >= 0)
            {
                if (stepsLeft-- <= 0)
                {
                    fault = HIPCSTONE_FAULT_WALK_BUDGET;
                    break;
                }
                if (listCount > kLeafList - 9) drain();
                int k      = uniform(sh.next[w][depth]);
                uint64_t m = 0;
                while (k < 8)
                {
                    m = uniform64(sh.mask[w][depth][k]);
                    if (m != 0) break;
                    ++k;
                }
                if (k >= 8)
                {
                    --depth;
                    continue;
                }
                int32_t node = uniform(sh.co[w][depth]) + k;
                if (lane == 0) sh.next[w][depth] = k + 1;
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

    if (valid) neighborsCount[i - firstId] = SlotMajor ? (myWpos - rowOff) >> 6 : myWpos - rowOff;
}

// the segmented search of the batched clouds (row-major rows, open boxes): see findNeighborsKernel
template<class T>
hipcstone_status findNeighborsSegmented(hipStream_t s, const T* x, const T* y, const T* z, const T* h,
                                        const typename BoxOf<T>::type& box, const int32_t* childOffsets,
                                        const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                        const T* sizes, int32_t numNodesCap, uint32_t ngmax, uint32_t* neighbors,
                                        uint32_t* neighborsCount, const uint32_t* waves, uint32_t numWaves,
                                        const int32_t* roots)
{
    if (numWaves == 0) return HIPCSTONE_OK;
    if (ngmax > (1u << 20)) return invalid("findNeighbors: ngmax above 2^20");
    uint32_t* faults = faultWord();
    if (!faults) return HIPCSTONE_HIP_ERROR;
    auto launch = [&](auto wv) {
        constexpr int W = decltype(wv)::value;
        unsigned blocks = ceilDiv(size_t(numWaves), W);
        blocks          = (blocks + kXcds - 1) / kXcds * kXcds;
        findNeighborsKernel<T, false, false, true, W><<<blocks, W * kWave, 0, s>>>(
            x, y, z, h, 0u, 0u, box, childOffsets, internalToLeaf, layout, centers, sizes, numNodesCap, 1.0f, ngmax,
            neighbors, neighborsCount, blocks, faults, waves, numWaves, roots);
    };
    switch (neighborWaves())
    {
    case 2: launch(std::integral_constant<int, 2>{}); break;
    case 8: launch(std::integral_constant<int, 8>{}); break;
    default: launch(std::integral_constant<int, 4>{}); break;
    }
    HC_CHECK_LAUNCH("findNeighborsKernel (segmented)");
    return HIPCSTONE_OK;
}

template hipcstone_status findNeighborsSegmented<float>(hipStream_t, const float*, const float*, const float*,
                                                        const float*, const BoxOf<float>::type&, const int32_t*,
                                                        const int32_t*, const uint32_t*, const float*, const float*,
                                                        int32_t, uint32_t, uint32_t*, uint32_t*, const uint32_t*,
                                                        uint32_t, const int32_t*);

template<class T>
static hipcstone_status findNeighbors(hipcstone_stream s, const T* x, const T* y, const T* z, const T* h,
                                      uint32_t firstId, uint32_t lastId, const typename BoxOf<T>::type* box,
                                      const int32_t* childOffsets, const int32_t* parents,
                                      const int32_t* internalToLeaf, const uint32_t* layout, const T* centers,
                                      const T* sizes, int32_t numNodes, float ext, uint32_t ngmax,
                                      uint32_t* neighbors, uint32_t* neighborsCount, bool slotMajor)
{
    if (lastId <= firstId) return HIPCSTONE_OK;
    if (!x || !y || !z || !h || !box || !childOffsets || !parents || !internalToLeaf || !layout || !centers ||
        !sizes || !neighborsCount)
        return invalid("findNeighbors: null pointer");
    if (ngmax > 0 && !neighbors) return invalid("findNeighbors: neighbors is null with ngmax > 0");
    if (ngmax > (1u << 20)) return invalid("findNeighbors: ngmax above 2^20");
    if (numNodes <= 0) return invalid("findNeighbors: numNodes must be positive");
    size_t groups   = (size_t(lastId - firstId) + kWave - 1) / kWave;
    uint32_t* faults = faultWord();
    if (!faults) return HIPCSTONE_HIP_ERROR;
    bool anyPbc     = box->bnd[0] == HIPCSTONE_PERIODIC || box->bnd[1] == HIPCSTONE_PERIODIC ||
                  box->bnd[2] == HIPCSTONE_PERIODIC;
    auto shaped = [&](auto pbc, auto slots, auto wv) {
        constexpr int W = decltype(wv)::value;
        unsigned blocks = ceilDiv(groups, W);
        blocks          = (blocks + kXcds - 1) / kXcds * kXcds; // whole rounds over the XCDs; extra groups exit
        findNeighborsKernel<T, decltype(pbc)::value, decltype(slots)::value, false, W>
            <<<blocks, W * kWave, 0, stream(s)>>>(x, y, z, h, firstId, lastId, *box, childOffsets, internalToLeaf,
                                                  layout, centers, sizes, numNodes, ext, ngmax, neighbors,
                                                  neighborsCount, blocks, faults);
    };
    auto launch = [&](auto pbc, auto slots) {
        switch (neighborWaves())
        {
        case 2: shaped(pbc, slots, std::integral_constant<int, 2>{}); break;
        case 8: shaped(pbc, slots, std::integral_constant<int, 8>{}); break;
        default: shaped(pbc, slots, std::integral_constant<int, 4>{}); break;
        }
    };
    if (anyPbc && slotMajor) launch(std::true_type{}, std::true_type{});
    else if (anyPbc) launch(std::true_type{}, std::false_type{});
    else if (slotMajor) launch(std::false_type{}, std::true_type{});
    else launch(std::false_type{}, std::false_type{});
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
                                layout, centers, sizes, num_nodes, ext, ngmax, neighbors, neighbors_count, false);     \
    }                                                                                                                  \
    hipcstone_status hipcstone_find_neighbors_slot_major_##RS(                                                         \
        hipcstone_stream s, const T* x, const T* y, const T* z, const T* h, uint32_t first_id, uint32_t last_id,      \
        const hipcstone_box_##RS* box, const int32_t* child_offsets, const int32_t* parents,                          \
        const int32_t* internal_to_leaf, const uint32_t* layout, const T* centers, const T* sizes, int32_t num_nodes,  \
        float ext, uint32_t ngmax, uint32_t* neighbors, uint32_t* neighbors_count)                                     \
    {                                                                                                                  \
        return findNeighbors<T>(s, x, y, z, h, first_id, last_id, box, child_offsets, parents, internal_to_leaf,      \
                                layout, centers, sizes, num_nodes, ext, ngmax, neighbors, neighbors_count, true);      \
    }

HIPCSTONE_DEF_NEIGHBORS(f32, float)
HIPCSTONE_DEF_NEIGHBORS(f64, double)

} // extern "C"
