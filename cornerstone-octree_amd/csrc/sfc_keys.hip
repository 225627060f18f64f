// a2/a3: SFC key encoding, keys[i] = sfc3D(x[i], y[i], z[i], box).
//
// Reference semantics: sfc/sfc.hpp:142-179 (sfc3D), :268-276 (computeSfcKeys), sfc/sfc_gpu.cu:24-33 (removeKey
// skip). One particle per lane, coalesced SoA loads; the per-particle Hilbert loop (21 levels for 64-bit keys)
// makes this kernel ALU-bound rather than HBM-bound.
#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

template<class K, int Curve, class T, bool HonorRemove>
__global__ __launch_bounds__(256) void computeSfcKeysKernel(const T* __restrict__ x, const T* __restrict__ y,
                                                            const T* __restrict__ z, K* __restrict__ keys, size_t n,
                                                            typename BoxOf<T>::type box)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (HonorRemove)
    {
        constexpr K rmKey = K(K(1) << (3 * maxLevel<K>()));
        if (keys[i] == rmKey) return;
    }
    keys[i] = sfc3D<K, Curve>(x[i], y[i], z[i], box);
}

template<class K, int Curve, class T>
static hipcstone_status computeSfcKeys(hipcstone_stream s, const T* x, const T* y, const T* z, K* keys, size_t n,
                                       const typename BoxOf<T>::type* box, int honorRemove)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!x || !y || !z || !keys || !box) return invalid("computeSfcKeys: null pointer");
    unsigned blocks = ceilDiv(n, 256);
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
