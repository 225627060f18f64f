// Synthetic-workload helpers for the benchmarks and tests: per-particle random displacement
// x_i += scale * h_i * u_i with u_i uniform in [-1, 1)^3 from a counter-based hash of (seed, step, i, dim),
// clamped into the box, and a Gaussian-mixture particle generator with h matched to the mixture density. Not part of the octree
// path itself; they produce the inputs that the path consumes.
#include <limits>

#include "common.hpp"
#include "runtime.hpp"
#include "synth.hpp"

namespace hipcstone
{

template<class T>
__global__ __launch_bounds__(256) void displaceKernel(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z,
                                                      const T* __restrict__ h, size_t n, T scale, uint64_t seed,
                                                      uint64_t step, typename BoxOf<T>::type box)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    displaceAt(x, y, z, h, i, uint64_t(i), scale, seed, step, box);
}

template<class T>
__global__ __launch_bounds__(256) void displaceNormalKernel(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z,
                                                            const T* __restrict__ h, size_t n, T sigma, uint64_t seed,
                                                            uint64_t step, typename BoxOf<T>::type box)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    displaceNormalAt(x, y, z, h, i, uint64_t(i), sigma, seed, step, box);
}

// value k of stream `stream` under `seed`, in [0, 1): the counter-based splitmix64 sequence of the host-side input
// generators (same bits for the same (seed, stream, k)).
__device__ inline double uniformAt(uint64_t seed, uint64_t stream, uint64_t k)
{
    uint64_t base = mix64(seed * 0x9E3779B97F4A7C15ull + stream);
    return double(mix64(base + k) >> 11) * (1.0 / 9007199254740992.0);
}

// Gaussian-mixture particles: `clusters` isotropic Gaussians (centres uniform in [0,1)^3 from streams 10..12, sigma
// log-uniform in [0.005, 0.05] from stream 13) plus a uniform background fraction (stream 14 selects, streams 20..22
// place); cluster choice from stream 15, Box-Muller normals from streams 16..19; coordinates clipped to [0, 1).
// h = 0.5 (target / (4/3 pi rho))^(1/3) with rho the mixture's density at the particle for n_total particles
// (background plus every cluster), so that h follows the local density as an SPH smoothing length does.
// Particle i is element offset + i of every stream. Up to kMaxClusters clusters.
constexpr int kMaxClusters = 256;

template<class T>
__global__ __launch_bounds__(256) void gaussianMixtureKernel(T* __restrict__ x, T* __restrict__ y, T* __restrict__ z,
                                                             T* __restrict__ h, size_t n, uint64_t offset,
                                                             uint64_t seed, uint32_t clusters, double background,
                                                             double target, double nTotal)
{
    __shared__ double cx[kMaxClusters], cy[kMaxClusters], cz[kMaxClusters], cg[kMaxClusters], cs[kMaxClusters],
        cn[kMaxClusters];
    for (uint32_t c = threadIdx.x; c < clusters; c += blockDim.x)
    {
        double sig = exp(log(0.005) + (log(0.05) - log(0.005)) * uniformAt(seed, 13, c));
        cx[c]      = uniformAt(seed, 10, c);
        cy[c]      = uniformAt(seed, 11, c);
        cz[c]      = uniformAt(seed, 12, c);
        cg[c]      = sig;
        cs[c]      = -0.5 / (sig * sig);
        cn[c]      = (1.0 - background) / double(clusters) * nTotal / (pow(2.0 * M_PI, 1.5) * sig * sig * sig);
    }
    __syncthreads();
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t k = offset + i;
    double p[3];
    if (uniformAt(seed, 14, k) < background)
    {
        for (int d = 0; d < 3; ++d)
            p[d] = uniformAt(seed, 20 + d, k);
    }
    else
    {
        uint32_t c = uint32_t(uniformAt(seed, 15, k) * clusters);
        c          = c < clusters - 1 ? c : clusters - 1;
        double sig = cg[c];
        double u1  = uniformAt(seed, 16, k), u2 = uniformAt(seed, 17, k);
        double u3  = uniformAt(seed, 18, k), u4 = uniformAt(seed, 19, k);
        u1         = u1 > 1e-300 ? u1 : 1e-300;
        u3         = u3 > 1e-300 ? u3 : 1e-300;
        double r1  = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
        double g[3] = {r1 * cos(2.0 * M_PI * u2), r1 * sin(2.0 * M_PI * u2), r2 * cos(2.0 * M_PI * u4)};
        double ctr[3] = {cx[c], cy[c], cz[c]};
        for (int d = 0; d < 3; ++d)
            p[d] = ctr[d] + sig * g[d];
    }
    const T top = T(1) - std::numeric_limits<T>::epsilon() / T(2); // largest T below 1
    T q[3];
    for (int d = 0; d < 3; ++d)
    {
        T v  = T(p[d] < 0.0 ? 0.0 : p[d]);
        q[d] = v < top ? v : top;
    }
    double rho = background * nTotal;
    for (uint32_t c = 0; c < clusters; ++c)
    {
        double dx = double(q[0]) - cx[c], dy = double(q[1]) - cy[c], dz = double(q[2]) - cz[c];
        rho += cn[c] * exp(cs[c] * (dx * dx + dy * dy + dz * dz));
    }
    double hh = 0.5 * cbrt(target / (4.0 / 3.0 * M_PI * rho));
    x[i]      = q[0];
    y[i]      = q[1];
    z[i]      = q[2];
    h[i]      = T(hh);
}

template<class T>
static hipcstone_status gaussianMixture(hipcstone_stream s, T* x, T* y, T* z, T* h, size_t n, uint64_t offset,
                                        uint64_t seed, uint32_t clusters, double background, double target,
                                        double nTotal)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!x || !y || !z || !h) return invalid("gaussianMixture: null pointer");
    if (clusters == 0 || clusters > uint32_t(kMaxClusters) || !(background >= 0.0 && background <= 1.0) ||
        !(nTotal > 0.0) || !(target > 0.0))
        return invalid("gaussianMixture: bad parameters");
    gaussianMixtureKernel<T><<<ceilDiv(n, 256), 256, 0, stream(s)>>>(x, y, z, h, n, offset, seed, clusters,
                                                                     background, target, nTotal);
    HC_CHECK_LAUNCH("gaussianMixtureKernel");
    return HIPCSTONE_OK;
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

hipcstone_status hipcstone_synth_displace_normal_f32(hipcstone_stream s, float* x, float* y, float* z,
                                                     const float* h, size_t n, float sigma, uint64_t seed,
                                                     uint64_t step, const hipcstone_box_f32* box)
{
    if (n == 0) return HIPCSTONE_OK;
    if (!x || !y || !z || !h || !box) return invalid("synth_displace_normal: null pointer");
    displaceNormalKernel<float><<<ceilDiv(n, 256), 256, 0, stream(s)>>>(x, y, z, h, n, sigma, seed, step, *box);
    HC_CHECK_LAUNCH("displaceNormalKernel");
    return HIPCSTONE_OK;
}

hipcstone_status hipcstone_synth_gaussian_mixture_f32(hipcstone_stream s, float* x, float* y, float* z, float* h,
                                                      size_t n, uint64_t offset, uint64_t seed, uint32_t clusters,
                                                      double background, double target_neighbors, double n_total)
{
    return gaussianMixture<float>(s, x, y, z, h, n, offset, seed, clusters, background, target_neighbors, n_total);
}

hipcstone_status hipcstone_synth_gaussian_mixture_f64(hipcstone_stream s, double* x, double* y, double* z, double* h,
                                                      size_t n, uint64_t offset, uint64_t seed, uint32_t clusters,
                                                      double background, double target_neighbors, double n_total)
{
    return gaussianMixture<double>(s, x, y, z, h, n, offset, seed, clusters, background, target_neighbors, n_total);
}

} // extern "C"
