// Synthetic-workload displacement shared by the plain and batched drift steps (not part of the octree path):
// particle i of a set displaced under (seed, step) moves by scale * h_i * u_i, u_i uniform in [-1, 1)^3 from a
// counter-based hash of (seed, step, i, dim), and is clamped into the box.
#pragma once

#include "common.hpp"

namespace hipcstone
{

template<class T, class Box>
__device__ inline void displaceAt(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z, const T* __restrict__ h,
                                  size_t at, uint64_t i, T scale, uint64_t seed, uint64_t step, const Box& box)
{
    uint64_t base = mix64(seed * 0x100000001B3ull + step);
    T* c[3]       = {x, y, z};
    T amp         = scale * h[at];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        uint64_t r = mix64(base + 3 * i + d);
        T u        = T(double(r >> 11) * (1.0 / 9007199254740992.0)) * T(2) - T(1);
        T v        = c[d][at] + amp * u;
        T lo = box.lim[2 * d], hi = box.lim[2 * d + 1];
        v        = v < lo ? lo : (v > hi ? hi : v);
        c[d][at] = v;
    }
}

} // namespace hipcstone
