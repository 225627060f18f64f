// Device-side building blocks shared by all hipcstone kernels (gfx950, wave64).
//
// SFC key arithmetic and geometry restated from the reference (sekelle/cornerstone-octree, include/cstone/):
// every function names the file:line whose semantics it keeps. All floating-point expressions keep the
// reference's evaluation order; the library is compiled with -ffp-contract=off so no FMA is formed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "hipcstone.h"
#include "hipcstone_nbwalk.hpp" // wave64 primitives, BoxOf, the neighbor walk

namespace hipcstone
{

template<class K>
struct KeyTraits;

// tree/definitions.h:29-60 (unusedBits, maxTreeLevel)
template<>
struct KeyTraits<uint32_t>
{
    static constexpr unsigned maxLevel = 10;
    static constexpr unsigned unused   = 2;
    static constexpr unsigned bits     = 32;
};

template<>
struct KeyTraits<uint64_t>
{
    static constexpr unsigned maxLevel = 21;
    static constexpr unsigned unused   = 1;
    static constexpr unsigned bits     = 64;
};

template<class K>
__host__ __device__ constexpr unsigned maxLevel()
{
    return KeyTraits<K>::maxLevel;
}

// primitives/clz.hpp:65-108: leading zeros, the bit width for an input of 0
__host__ __device__ inline int clzKey(uint32_t x) { return x == 0 ? 32 : __builtin_clz(x); }
__host__ __device__ inline int clzKey(uint64_t x) { return x == 0 ? 64 : __builtin_clzll(x); }

// sfc/common.hpp:109-116
template<class K>
__host__ __device__ inline K nodeRange(unsigned level)
{
    return K(K(1) << (3u * (maxLevel<K>() - level)));
}

// sfc/common.hpp:157-161
template<class K>
__host__ __device__ inline unsigned treeLevel(K range)
{
    return unsigned(clzKey(K(range - 1)) - int(KeyTraits<K>::unused)) / 3;
}

// sfc/common.hpp:119-125
template<class K>
__host__ __device__ inline unsigned log8ceil(K n)
{
    if (n == 0) { return 0; }
    unsigned lz = clzKey(K(n - 1));
    return maxLevel<K>() - (lz - KeyTraits<K>::unused) / 3;
}

// sfc/common.hpp:145-148
template<class K>
__host__ __device__ inline int commonPrefix(K a, K b)
{
    return clzKey(K(a ^ b)) - int(KeyTraits<K>::unused);
}

// sfc/common.hpp:173-180 (Warren-Salmon placeholder bit)
template<class K>
__host__ __device__ inline K encodePlaceholderBit(K code, int prefixLength)
{
    int nShifts = 3 * int(maxLevel<K>()) - prefixLength;
    return (K(1) << prefixLength) | (code >> nShifts);
}

// sfc/common.hpp:192-195
template<class K>
__host__ __device__ inline unsigned decodePrefixLength(K code)
{
    return 8 * sizeof(K) - 1 - clzKey(code);
}

// sfc/common.hpp:206-213
template<class K>
__host__ __device__ inline K decodePlaceholderBit(K code)
{
    int prefixLength = int(decodePrefixLength(code));
    K ret            = code ^ (K(1) << prefixLength);
    return ret << (3 * int(maxLevel<K>()) - prefixLength);
}

// sfc/common.hpp:299-302
template<class K>
__host__ __device__ inline unsigned octalDigit(K code, unsigned position)
{
    return unsigned((code >> (3u * (maxLevel<K>() - position))) & 7u);
}

// tree/octree.hpp:55-65 with sfc/common.hpp:318-322 digitWeight: the weight that places the internal node created
// between two adjacent leaves at its slot (tid + weight) / 7 of the unsorted internal-node layout
template<class K>
__host__ __device__ inline int32_t binaryKeyWeight(K key, unsigned level)
{
    int32_t ret = 0;
    for (unsigned l = 1; l <= level + 1; ++l)
    {
        int digit       = int(octalDigit(key, l));
        int fourGeqMask = -int(digit >= 4);
        ret += ((7 - digit) & fourGeqMask) - (digit & ~fourGeqMask);
    }
    return ret;
}

// splitmix64 finalizer: the counter-based hash of the synthetic input generators (displacements, mixtures)
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// sfc/common.hpp:326-331
template<class K>
__host__ __device__ inline K enclosingBoxCode(K key, unsigned level)
{
    return key & ~K(nodeRange<K>(level) - 1);
}

// sfc/morton.hpp:35-44 / :60-70 expandBits
__host__ __device__ inline uint32_t expandBits(uint32_t v)
{
    v &= 0x000003ffu;
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__host__ __device__ inline uint64_t expandBits(uint64_t v)
{
    uint64_t x = v & 0x1fffffu;
    x          = (x | x << 32u) & 0x001f00000000ffffull;
    x          = (x | x << 16u) & 0x001f0000ff0000ffull;
    x          = (x | x << 8u) & 0x100f00f00f00f00full;
    x          = (x | x << 4u) & 0x10c30c30c30c30c3ull;
    x          = (x | x << 2u) & 0x1249249249249249ull;
    return x;
}

// sfc/morton.hpp:95-108
template<class K>
__host__ __device__ inline K iMorton(unsigned ix, unsigned iy, unsigned iz)
{
    K xx = expandBits(K(ix));
    K yy = expandBits(K(iy));
    K zz = expandBits(K(iz));
    return xx * 4 + yy * 2 + zz;
}

// sfc/morton.hpp:48-56 / :75-85 compactBits
__host__ __device__ inline unsigned compactBits(uint32_t v)
{
    v &= 0x09249249u;
    v = (v ^ (v >> 2u)) & 0x030c30c3u;
    v = (v ^ (v >> 4u)) & 0x0300f00fu;
    v = (v ^ (v >> 8u)) & 0xff0000ffu;
    v = (v ^ (v >> 16u)) & 0x000003ffu;
    return v;
}

__host__ __device__ inline unsigned compactBits(uint64_t v)
{
    v &= 0x1249249249249249ull;
    v = (v ^ (v >> 2u)) & 0x10c30c30c30c30c3ull;
    v = (v ^ (v >> 4u)) & 0x100f00f00f00f00full;
    v = (v ^ (v >> 8u)) & 0x001f0000ff0000ffull;
    v = (v ^ (v >> 16u)) & 0x001f00000000ffffull;
    v = (v ^ (v >> 32u)) & 0x00000000001fffffull;
    return unsigned(v);
}

// Morton octant -> Hilbert digit, the table {0,1,3,2,7,6,4,5} of sfc/hilbert.hpp:34,52 packed into nibbles
constexpr uint32_t kMortonToHilbertNibbles = 0x54672310u;

// sfc/hilbert.hpp:43-94 iHilbert: per level, emit the Hilbert digit of the Morton octant and rotate/reflect
// the remaining coordinate bits. Branch-free form of the same transformation.
template<class K>
__host__ __device__ inline K iHilbert(unsigned px, unsigned py, unsigned pz)
{
    K key = 0;
#pragma unroll
    for (int level = int(maxLevel<K>()) - 1; level >= 0; --level)
    {
        unsigned xi = (px >> level) & 1u;
        unsigned yi = (py >> level) & 1u;
        unsigned zi = (pz >> level) & 1u;

        unsigned octant = (xi << 2) | (yi << 1) | zi;
        key             = (key << 3) + ((kMortonToHilbertNibbles >> (4 * octant)) & 7u);

        px ^= -(xi & ((!yi) | zi));
        py ^= -((xi & (yi | zi)) | (yi & (!zi)));
        pz ^= -((xi & (!yi) & (!zi)) | (yi & (!zi)));

        // zi: cyclic rotation (x,y,z) <- (y,z,x); else !yi: swap x and z
        unsigned nx = zi ? py : (yi ? px : pz);
        unsigned ny = zi ? pz : py;
        unsigned nz = zi ? px : (yi ? pz : px);
        px          = nx;
        py          = ny;
        pz          = nz;
    }
    return key;
}

// sfc/hilbert.hpp:130-172 decodeHilbert
template<class K>
__host__ __device__ inline void decodeHilbert(K key, unsigned& ox, unsigned& oy, unsigned& oz)
{
    unsigned px = 0, py = 0, pz = 0;
#pragma unroll
    for (unsigned level = 0; level < maxLevel<K>(); ++level)
    {
        unsigned octant   = unsigned((key >> (3 * level)) & 7u);
        const unsigned xi = octant >> 2u;
        const unsigned yi = (octant >> 1u) & 1u;
        const unsigned zi = octant & 1u;

        // yi^zi: (x,y,z) <- (z,x,y); else 000 or 111: swap x and z
        bool rot      = (yi ^ zi) != 0;
        bool swp      = !rot && ((!xi & !yi & !zi) || (xi & yi & zi));
        unsigned nx   = rot ? pz : (swp ? pz : px);
        unsigned ny   = rot ? px : py;
        unsigned nz   = rot ? py : (swp ? px : pz);
        px            = nx;
        py            = ny;
        pz            = nz;
        unsigned mask = (1u << level) - 1;
        px ^= mask & (-(xi & (yi | zi)));
        py ^= mask & (-((xi & ((!yi) | (!zi))) | ((!xi) & yi & zi)));
        pz ^= mask & (-((xi & (!yi) & (!zi)) | (yi & zi)));

        px |= (xi << level);
        py |= ((xi ^ yi) << level);
        pz |= ((yi ^ zi) << level);
    }
    ox = px;
    oy = py;
    oz = pz;
}

// sfc/hilbert.hpp:260-275 hilbertIBox: integer box {xmin,xmax,ymin,ymax,zmin,zmax} of the node at keyStart/level
template<class K>
__host__ __device__ inline void hilbertIBox(K keyStart, unsigned level, int ibox[6])
{
    unsigned cubeLength = (1u << maxLevel<K>()) >> level;
    unsigned mask       = ~(cubeLength - 1);
    unsigned ix, iy, iz;
    decodeHilbert(keyStart, ix, iy, iz);
    ix &= mask;
    iy &= mask;
    iz &= mask;
    ibox[0] = int(ix);
    ibox[1] = int(ix + cubeLength);
    ibox[2] = int(iy);
    ibox[3] = int(iy + cubeLength);
    ibox[4] = int(iz);
    ibox[5] = int(iz + cubeLength);
}

// sfc/common.hpp:343-349 smallestCommonBox
template<class K>
__host__ __device__ inline void smallestCommonBox(K a, K b, K& start, K& end)
{
    unsigned commonLevel = unsigned(commonPrefix(a, b)) / 3;
    start                = enclosingBoxCode(a, commonLevel);
    end                  = start + nodeRange<K>(commonLevel);
}

// sfc/sfc.hpp:142-179 sfc3D: ix = int(floor(x*mx) - xmin*mx) with mx = 2^L * ilx, clamped to 2^L - 1
template<class K, class T, class Box>
__host__ __device__ inline void sfcGridCoords(T x, T y, T z, const Box& box, int& ix, int& iy, int& iz)
{
    constexpr unsigned cubeLength = 1u << maxLevel<K>();
    constexpr int mcoord          = int(cubeLength - 1);
    T mx                          = T(cubeLength) * box.ilen[0];
    T my                          = T(cubeLength) * box.ilen[1];
    T mz                          = T(cubeLength) * box.ilen[2];
    ix                            = int(floor(x * mx) - box.lim[0] * mx);
    iy                            = int(floor(y * my) - box.lim[2] * my);
    iz                            = int(floor(z * mz) - box.lim[4] * mz);
    ix                            = ix < mcoord ? ix : mcoord;
    iy                            = iy < mcoord ? iy : mcoord;
    iz                            = iz < mcoord ? iz : mcoord;
}

template<class K, int Curve, class T, class Box>
__host__ __device__ inline K sfc3D(T x, T y, T z, const Box& box)
{
    int ix, iy, iz;
    sfcGridCoords<K>(x, y, z, box, ix, iy, iz);
    if constexpr (Curve == HIPCSTONE_HILBERT) { return iHilbert<K>(unsigned(ix), unsigned(iy), unsigned(iz)); }
    else { return iMorton<K>(unsigned(ix), unsigned(iy), unsigned(iz)); }
}

// ---- wave64 primitives (the rest are in hipcstone_nbwalk.hpp) ----

template<class T>
__device__ inline T readLane(T v, int lane);

template<>
__device__ inline int readLane(int v, int lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}
template<>
__device__ inline unsigned readLane(unsigned v, int lane)
{
    return unsigned(__builtin_amdgcn_readlane(int(v), lane));
}
template<>
__device__ inline float readLane(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template<>
__device__ inline double readLane(double v, int lane)
{
    int64_t b = __double_as_longlong(v);
    int lo    = __builtin_amdgcn_readlane(int(uint32_t(b)), lane);
    int hi    = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(b) >> 32)), lane);
    return __longlong_as_double(int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo)));
}

inline unsigned ceilDiv(size_t a, size_t b) { return unsigned((a + b - 1) / b); }

} // namespace hipcstone
