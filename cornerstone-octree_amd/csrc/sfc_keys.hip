// a2/a3: SFC key encoding, keys[i] = sfc3D(x[i], y[i], z[i], box).
//
// Reference semantics: sfc/sfc.hpp:142-179 (sfc3D), :268-276 (computeSfcKeys), sfc/sfc_gpu.cu:24-33 (removeKey
// skip). One particle per lane, coalesced SoA loads.
//
// Hilbert keys come from a state machine equivalent to iHilbert (sfc/hilbert.hpp:43-94): iHilbert's per-level update
// flips and permutes the remaining coordinate words, so everything it does below a level is fixed by a signed axis
// permutation (the state: 24 of them, the rotations of the cube). Tables built at compile time from that update map (state, octant of
// the raw coordinates) to the Hilbert digit and the next state, two levels per lookup (64 entries per state, in LDS);
// the raw octants of all levels are the octal digits of the Morton code. This replaces ~20 operations per level by one
// lookup per two levels and gives the same keys.
#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

// signed axis permutation: transformed coordinate d = raw coordinate src(d), bitwise inverted if inv(d)
struct HilbertState
{
    uint8_t src[3];
    uint8_t inv[3];
    constexpr bool operator==(const HilbertState& o) const
    {
        return src[0] == o.src[0] && src[1] == o.src[1] && src[2] == o.src[2] && inv[0] == o.inv[0] &&
               inv[1] == o.inv[1] && inv[2] == o.inv[2];
    }
};

// one level of iHilbert on the transformed bits (xi, yi, zi) of the raw octant r: the digit and the next state
constexpr HilbertState hilbertStep(const HilbertState& s, unsigned r, unsigned& digit)
{
    const unsigned raw[3] = {(r >> 2) & 1u, (r >> 1) & 1u, r & 1u};
    const unsigned xi = raw[s.src[0]] ^ s.inv[0], yi = raw[s.src[1]] ^ s.inv[1], zi = raw[s.src[2]] ^ s.inv[2];
    digit             = (kMortonToHilbertNibbles >> (4 * ((xi << 2) | (yi << 1) | zi))) & 7u;
    // the word flips of iHilbert
    const uint8_t inv[3] = {uint8_t(s.inv[0] ^ (xi & ((!yi) | zi))),
                            uint8_t(s.inv[1] ^ ((xi & (yi | zi)) | (yi & (!zi)))),
                            uint8_t(s.inv[2] ^ ((xi & (!yi) & (!zi)) | (yi & (!zi))))};
    // then its word permutation: zi: (x, y, z) <- (y, z, x); else !yi: swap x and z
    const int px = zi ? 1 : (yi ? 0 : 2);
    const int py = zi ? 2 : 1;
    const int pz = zi ? 0 : (yi ? 2 : 0);
    return HilbertState{{s.src[px], s.src[py], s.src[pz]}, {inv[px], inv[py], inv[pz]}};
}

constexpr int kMaxHilbertStates = 24;

struct HilbertTables
{
    int numStates;
    uint16_t one[kMaxHilbertStates * 8];  // digit | next state << 3
    uint16_t two[kMaxHilbertStates * 64]; // digit pair (6 bits) | next state << 6
};

constexpr HilbertTables makeHilbertTables()
{
    HilbertTables t{};
    HilbertState states[kMaxHilbertStates]{};
    states[0]  = HilbertState{{0, 1, 2}, {0, 0, 0}};
    int n      = 1;
    auto index = [&](const HilbertState& h) {
        for (int i = 0; i < n; ++i)
            if (states[i] == h) return i;
        states[n] = h;
        return n++;
    };
    for (int i = 0; i < n; ++i) // n grows while states are discovered
    {
        for (unsigned r = 0; r < 8; ++r)
        {
            unsigned d = 0;
            const int next = index(hilbertStep(states[i], r, d));
            t.one[i * 8 + r] = uint16_t(d | (unsigned(next) << 3));
        }
    }
    for (int i = 0; i < n; ++i)
    {
        for (unsigned r1 = 0; r1 < 8; ++r1)
        {
            const unsigned e1 = t.one[i * 8 + r1];
            const unsigned s1 = e1 >> 3;
            for (unsigned r2 = 0; r2 < 8; ++r2)
            {
                const unsigned e2           = t.one[s1 * 8 + r2];
                t.two[i * 64 + r1 * 8 + r2] = uint16_t(((e1 & 7u) << 3) | (e2 & 7u) | ((e2 >> 3) << 6));
            }
        }
    }
    t.numStates = n;
    return t;
}

constexpr HilbertTables kHilbertTables = makeHilbertTables();
static_assert(kHilbertTables.numStates <= kMaxHilbertStates, "Hilbert state count");

__constant__ HilbertTables kHilbertTablesDevice = kHilbertTables;

// Hilbert key of the grid point (ix, iy, iz) from the tables: the octal digits of the Morton code are the raw
// octants, from the top level down
template<class K>
__host__ __device__ inline K hilbertFromTables(unsigned ix, unsigned iy, unsigned iz, const uint16_t* one, const uint16_t* two)
{
    constexpr int L = int(maxLevel<K>());
    const K m       = iMorton<K>(ix, iy, iz);
    K key           = 0;
    unsigned st     = 0;
    int level       = L;
    if constexpr (L % 2 == 1)
    {
        const unsigned e = one[(st << 3) | unsigned(m >> (3 * (L - 1)))];
        key              = K(e & 7u);
        st               = e >> 3;
        --level;
    }
#pragma unroll
    for (; level >= 2; level -= 2)
    {
        const unsigned e = two[(st << 6) | unsigned((m >> (3 * (level - 2))) & 63u)];
        key              = (key << 6) | K(e & 63u);
        st               = e >> 6;
    }
    return key;
}

// grid-stride over the particles (so that each block's copy of the Hilbert tables serves many keys)
template<class K, int Curve, class T, bool HonorRemove>
__global__ __launch_bounds__(256) void computeSfcKeysKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                            const T* __restrict__ z, K* __restrict__ keys, size_t n,
                                                            typename BoxOf<T>::type box)
{
    __shared__ uint16_t one[kMaxHilbertStates * 8];
    __shared__ uint16_t two[kMaxHilbertStates * 64];
    if constexpr (Curve == HIPCSTONE_HILBERT)
    {
        const int ns = kHilbertTables.numStates;
        for (int t = threadIdx.x; t < ns * 64; t += blockDim.x)
            two[t] = kHilbertTablesDevice.two[t];
        for (int t = threadIdx.x; t < ns * 8; t += blockDim.x)
            one[t] = kHilbertTablesDevice.one[t];
        __syncthreads();
    }
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
    {
        if constexpr (HonorRemove)
        {
            constexpr K rmKey = K(K(1) << (3 * maxLevel<K>()));
            if (keys[i] == rmKey) continue;
        }
        if constexpr (Curve == HIPCSTONE_HILBERT)
        {
            int ix, iy, iz;
            sfcGridCoords<K>(x[i], y[i], z[i], box, ix, iy, iz);
            keys[i] = hilbertFromTables<K>(unsigned(ix), unsigned(iy), unsigned(iz), one, two);
        }
        else { keys[i] = sfc3D<K, Curve>(x[i], y[i], z[i], box); }
    }
}

template<class K, int Curve, class T>
static hipcstone_status computeSfcKeys(hipcstone_stream s, const T* x, const T* y, const T* z, K* keys, size_t n,
                                       const typename BoxOf<T>::type* box, int honorRemove)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!x || !y || !z || !keys || !box) return invalid("computeSfcKeys: null pointer");
    unsigned blocks = ceilDiv(n, 256);
    // Hilbert: 16 blocks per CU, each block's table copy then serves several passes of the grid-stride loop
    if constexpr (Curve == HIPCSTONE_HILBERT) { blocks = blocks < 4096u ? blocks : 4096u; }
    if (honorRemove)
        computeSfcKeysKernel<K, Curve, T, true><<<blocks, 256, 0, stream(s)>>>(x, y, z, keys, n, *box);
    else
        computeSfcKeysKernel<K, Curve, T, false><<<blocks, 256, 0, stream(s)>>>(x, y, z, keys, n, *box);
    HC_CHECK_LAUNCH("computeSfcKeysKernel");
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

#define HIPCSTONE_DEF_KEYS(CURVE, CURVEID, KS, K, RS, T)                                                               \
    extern "C" hipcstone_status hipcstone_compute_sfc_keys_##CURVE##_##KS##_##RS(                                      \
        hipcstone_stream s, const T* x, const T* y, const T* z, K* keys, size_t n, const hipcstone_box_##RS* box,     \
        int honor_remove_key)                                                                                          \
    {                                                                                                                  \
        return computeSfcKeys<K, CURVEID, T>(s, x, y, z, keys, n, box, honor_remove_key);                             \
    }

HIPCSTONE_DEF_KEYS(hilbert, HIPCSTONE_HILBERT, u32, uint32_t, f32, float)
HIPCSTONE_DEF_KEYS(hilbert, HIPCSTONE_HILBERT, u32, uint32_t, f64, double)
HIPCSTONE_DEF_KEYS(hilbert, HIPCSTONE_HILBERT, u64, uint64_t, f32, float)
HIPCSTONE_DEF_KEYS(hilbert, HIPCSTONE_HILBERT, u64, uint64_t, f64, double)
HIPCSTONE_DEF_KEYS(morton, HIPCSTONE_MORTON, u32, uint32_t, f32, float)
HIPCSTONE_DEF_KEYS(morton, HIPCSTONE_MORTON, u32, uint32_t, f64, double)
HIPCSTONE_DEF_KEYS(morton, HIPCSTONE_MORTON, u64, uint64_t, f32, float)
HIPCSTONE_DEF_KEYS(morton, HIPCSTONE_MORTON, u64, uint64_t, f64, double)
