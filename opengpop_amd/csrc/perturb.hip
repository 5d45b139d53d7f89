// Monte-Carlo initial states around one orbit state: PerturbationGenerator
// (perturbation_generator.cu:21-79). Every satellite starts at the epoch position with the epoch
// velocity plus strength * N(0, 1) per component.
//
// The reference draws its normals with cuRAND (seed 1234, one subsequence per satellite). This
// generator uses Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011), keyed by the seed with the satellite index in the counter, and the
// Box-Muller transform: each satellite's numbers depend only on (seed, index), whatever the
// launch shape, so any subset of a population can be regenerated on its own.
#include "kernels.hpp"

#include <math.h>

namespace ogp {
namespace {

struct Words {
    uint32_t w[4];
};

// Ten Philox rounds: two 32 x 32 -> 64-bit products per round, the key advanced by the Weyl
// increments between rounds.
__device__ __forceinline__ Words philox4x32_10(Words c, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
    constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t p0 = (uint64_t)kM0 * c.w[0];
        const uint64_t p1 = (uint64_t)kM1 * c.w[2];
        const Words n { { (uint32_t)(p1 >> 32) ^ c.w[1] ^ k0, (uint32_t)p1,
            (uint32_t)(p0 >> 32) ^ c.w[3] ^ k1, (uint32_t)p0 } };
        c = n;
        k0 += kW0;
        k1 += kW1;
    }
    return c;
}

// 53 random bits of (hi, lo) as a double in the open interval (0, 1).
__device__ __forceinline__ double open_unit(uint32_t hi, uint32_t lo)
{
    const uint64_t bits = ((uint64_t)hi << 21) | (uint64_t)(lo >> 11);
    return ((double)bits + 0.5) * 0x1.0p-53;
}

// The three normals of satellite i: Box-Muller on the first counter block gives two, the cosine
// branch of the second block the third.
// The angle 2 pi u goes through sincospi/cospi of 2u (exact argument reduction, both functions of
// one angle together).
__device__ __forceinline__ void normals3(uint64_t i, uint32_t k0, uint32_t k1, double z[3])
{
    const uint32_t lo = (uint32_t)i, hi = (uint32_t)(i >> 32);
    const Words a = philox4x32_10(Words { { lo, hi, 0u, 0u } }, k0, k1);
    const Words b = philox4x32_10(Words { { lo, hi, 1u, 0u } }, k0, k1);
    const double ra = sqrt(-2.0 * log(open_unit(a.w[0], a.w[1])));
    const double rb = sqrt(-2.0 * log(open_unit(b.w[0], b.w[1])));
    double sa, ca;
    sincospi(2.0 * open_unit(a.w[2], a.w[3]), &sa, &ca);
    z[0] = ra * ca;
    z[1] = ra * sa;
    z[2] = rb * cospi(2.0 * open_unit(b.w[2], b.w[3]));
}

// One thread per satellite; thread n (when `append`) writes the appended object. AoS output
// ([i][3]) when stride == 0, else SoA (component c of satellite i at [c * stride + i]).
__global__ void __launch_bounds__(256) perturb_kernel(PerturbArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > a.n || (i == a.n && !a.append))
        return;
    double r[3], v[3];
    if (i == a.n) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c] = a.obj_r[c];
            v[c] = a.obj_v[c];
        }
    } else {
        double z[3];
        normals3((uint64_t)i, a.k0, a.k1, z);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            r[c] = a.r0[c];
            v[c] = a.v0[c] + z[c] * a.strength;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int64_t at = a.stride > 0 ? c * a.stride + i : 3 * i + c;
        a.r[at] = r[c];
        a.v[at] = v[c];
    }
}

} // namespace

void launch_perturb(hipStream_t st, const PerturbArgs& a)
{
    const int64_t count = a.n + (a.append ? 1 : 0);
    if (count <= 0)
        return;
    const dim3 grid((unsigned)((count + 255) / 256));
    hipLaunchKernelGGL(perturb_kernel, grid, dim3(256), 0, st, a);
}

} // namespace ogp
