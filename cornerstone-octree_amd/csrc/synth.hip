// Synthetic-workload helper for the multi-step sync benchmarks and tests: per-particle random displacement
// x_i += scale * h_i * u_i with u_i uniform in [-1, 1)^3 from a counter-based hash of (seed, step, i, dim),
// clamped into the box. Not part of the octree path itself; it produces the moving input that the warm-start
// steps consume.
#include "common.hpp"
#include "runtime.hpp"

namespace hipcstone
{

__device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template<class T>
__global__ __launch_bounds__(256) void displaceKernel(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z,
                                                      const T* __restrict__ h, size_t n, T scale, uint64_t seed,
                                                      uint64_t step, typename BoxOf<T>::type box)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t base = mix64(seed * 0x100000001B3ull + step);
    T* c[3]       = {x, y, z};
    T amp         = scale * h[i];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        uint64_t r = mix64(base + 3 * uint64_t(i) + d);
        T u        = T(double(r >> 11) * (1.0 / 9007199254740992.0)) * T(2) - T(1);
        T v        = c[d][i] + amp * u;
        T lo = box.lim[2 * d], hi = box.lim[2 * d + 1];
        v       = v < lo ? lo : (v > hi ? hi : v);
        c[d][i] = v;
    }
}

template<class T>
static hipcstone_status displace(hipcstone_stream s, T* x, T* y, T* z, const T* h, size_t n, T scale, uint64_t seed,
                                 uint64_t step, const typename BoxOf<T>::type* box)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!x || !y || !z || !h || !box) return invalid("displace: null pointer");
    displaceKernel<T><<<ceilDiv(n, 256), 256, 0, stream(s)>>>(x, y, z, h, n, scale, seed, step, *box);
    HC_CHECK_LAUNCH("displaceKernel");
    return HIPCSTONE_OK;
}

} // namespace hipcstone

using namespace hipcstone;

extern "C"
{

hipcstone_status hipcstone_synth_displace_f32(hipcstone_stream s, float* x, float* y, float* z, const float* h,
                                              size_t n, float scale, uint64_t seed, uint64_t step,
                                              const hipcstone_box_f32* box)
{
    return displace<float>(s, x, y, z, h, n, scale, seed, step, box);
}

hipcstone_status hipcstone_synth_displace_f64(hipcstone_stream s, double* x, double* y, double* z, const double* h,
                                              size_t n, double scale, uint64_t seed, uint64_t step,
                                              const hipcstone_box_f64* box)
{
    return displace<double>(s, x, y, z, h, n, scale, seed, step, box);
}

} // extern "C"
