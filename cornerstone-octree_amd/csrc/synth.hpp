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

// the same with a normal displacement: particle i moves by sigma * h_i * g_i, g_i ~ N(0, 1)^3 from Box-Muller pairs
// of the counter-based uniforms (dims 0, 1 from one pair, dim 2 from the next)
template<class T, class Box>
__device__ inline void displaceNormalAt(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z,
                                        const T* __restrict__ h, size_t at, uint64_t i, T sigma, uint64_t seed,
                                        uint64_t step, const Box& box)
{
    uint64_t base = mix64(seed * 0x100000001B3ull + step + 0x5DEECE66Dull);
    double u[4];
#pragma unroll
    for (int d = 0; d < 4; ++d)
        u[d] = double(mix64(base + 4 * i + d) >> 11) * (1.0 / 9007199254740992.0);
    double r0 = sqrt(-2.0 * log(1.0 - u[0])), r1 = sqrt(-2.0 * log(1.0 - u[2]));
    double g[3] = {r0 * cos(6.283185307179586 * u[1]), r0 * sin(6.283185307179586 * u[1]),
                   r1 * cos(6.283185307179586 * u[3])};
    T* c[3] = {x, y, z};
    T amp   = sigma * h[at];
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
        T v = c[d][at] + amp * T(g[d]);
        T lo = box.lim[2 * d], hi = box.lim[2 * d + 1];
        v        = v < lo ? lo : (v > hi ? hi : v);
        c[d][at] = v;
    }
}

} // namespace hipcstone
