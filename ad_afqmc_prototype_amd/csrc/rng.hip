// Device random stream of the step: keys split and normals / uniforms drawn
// on the GPU, so a step or a whole block has no host round trip (and can be
// captured in a graph).  The stream follows the reference's calls
// (prop/afqmc.py:32,86-88, prop/blocks.py:81-82) under jax's original
// (non-partitionable) Threefry scheme, config.py:27:
//
//   split(key)       = threefry(key, [0,1,2,3]) as (2,2): new key row 0, subkey row 1
//   (the caller splits, then draws from the subkey, as the reference does)
//   bits64[i] of n   = (A[i] << 32) | B[i],  (A[i], B[i]) = threefry(key, (i, n + i))
//   uniform(lo, hi)  = max(lo, (bitcast((bits >> 12) | bits(1.0)) - 1) * (hi - lo) + lo)
//   normal           = sqrt(2) erfinv(uniform(nextafter(-1, 0), 1))
//
// Threefry-2x32 with 20 rounds: Salmon, Moraes, Dror & Shaw, SC 2011.  erfinv is
// computed here by Halley iterations on erf / erfc (Winitzki's closed form as
// the start), so normals agree with another erfinv implementation to a few
// ulp, not bit for bit.
#include <cmath>
#include <cstdint>

#include "afqmc_hip.h"
#include "common.h"
#include "device_util.h"

namespace {

// sqrt(2) erfinv(u) for u in (-1, 1): closed-form start (|rel err| < 4e-3)
// and two Halley steps in fp32 (~1e-7 -> below fp32 rounding), then one
// Halley step in fp64 (cubic: to fp64 rounding), each on erf(y) = |u|
// (|u| <= 0.5) or erfc(y) = 1 - |u| (tails, where 1 - |u| is exact and erfc
// keeps full relative precision; 1 - |u| >= 2^-53 stays a normal fp32 value).
// Against scipy's erfinv on 2.4M uniforms, tails to 1 - 2^-53 included: 9.6e-16
// relative at most, as with three fp64 steps.
__device__ double sqrt2_erfinv(double u) {
    const double a = fabs(u);
    const double om = 1.0 - a;  // exact
    const float af = (float)a, omf = (float)om;
    const float t = logf(omf * (1.0f + af));
    const float c = (float)(2.0 / (M_PI * 0.147)) + 0.5f * t;
    float yf = sqrtf(sqrtf(c * c - t / 0.147f) - c);
    const float two_over_sqrt_pi_f = 1.12837917f;
    for (int it = 0; it < 2; ++it) {
        const float r = a <= 0.5 ? erff(yf) - af : omf - erfcf(yf);
        const float step = r / (two_over_sqrt_pi_f * expf(-yf * yf));
        yf -= step / (1.0f + yf * step);
    }
    double y = (double)yf;
    const double two_over_sqrt_pi = 1.1283791670955126;
    const double r = a <= 0.5 ? erf(y) - a : om - erfc(y);
    const double step = r / (two_over_sqrt_pi * exp(-y * y));  // erf'(y)
    y -= step / (1.0 + y * step);  // Halley: f'' / f' = -2 y
    return copysign(1.4142135623730951 * y, u);
}
__global__ void split_kernel(const uint32_t* __restrict__ key, uint32_t* __restrict__ new_key,
                             uint32_t* __restrict__ subkey) {
    uint32_t nk[2], sk[2];
    split_key(key, nk, sk);
    new_key[0] = nk[0];
    new_key[1] = nk[1];
    subkey[0] = sk[0];
    subkey[1] = sk[1];
}

// value i of normal(key, n values)
__device__ __forceinline__ double normal_value(uint32_t k0, uint32_t k1, long long n, long long i) {
    const double lo = -1.0 + 0x1p-53;  // nextafter(-1, 0)
    const double scale = 1.0 - lo;     // rounds to 2
    uint32_t a = (uint32_t)i, b = (uint32_t)(n + i);
    threefry2x32(k0, k1, a, b);
    const double f = unit_from_bits(a, b);
    return sqrt2_erfinv(fmax(lo, f * scale + lo));
}

__global__ void normal_kernel(const uint32_t* __restrict__ key, long long n, double* __restrict__ out) {
    const uint32_t k0 = key[0], k1 = key[1];
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = normal_value(k0, k1, n, i);
}

// One workgroup per walker row of every system: key_out, sub = split(key_in);
// fields = normal(sub, (nw, naux)).
__global__ __launch_bounds__(256) void fields_grouped_kernel(int count,
                                                             const afq_system_step* __restrict__ S) {
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    const int w = b - d.walker0;
    uint32_t nk[2], sk[2];
    split_key((const uint32_t*)d.key_in, nk, sk);
    if (w == 0 && threadIdx.x == 0) {
        d.key_out[0] = nk[0];
        d.key_out[1] = nk[1];
    }
    const long long n = (long long)d.nw * d.naux;
    const long long row = (long long)w * d.naux;
    for (int g = threadIdx.x; g < d.naux; g += 256) d.fields[row + g] = normal_value(sk[0], sk[1], n, row + g);
}

__global__ void uniform_kernel(const uint32_t* __restrict__ key, double* __restrict__ out) {
    out[0] = uniform_value(key[0], key[1]);
}

__global__ void threefry_kernel(uint32_t k0, uint32_t k1, long long n, const uint32_t* __restrict__ c0,
                                const uint32_t* __restrict__ c1, uint32_t* __restrict__ o0,
                                uint32_t* __restrict__ o1) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        uint32_t a = c0[i], b = c1[i];
        threefry2x32(k0, k1, a, b);
        o0[i] = a;
        o1[i] = b;
    }
}

unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

}  // namespace

extern "C" int afq_rng_split(const unsigned* key, unsigned* new_key, unsigned* subkey, void* stream) {
    AFQ_REQUIRE(key && new_key && subkey, "afq_rng_split: null pointer");
    AFQ_REQUIRE(key != new_key && key != subkey && new_key != subkey, "afq_rng_split: aliased keys");
    hipLaunchKernelGGL(split_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, (const uint32_t*)key,
                       (uint32_t*)new_key, (uint32_t*)subkey);
    return afq::check_launch("afq_rng_split");
}

extern "C" int afq_rng_normal(const unsigned* key, long long n, double* out, void* stream) {
    AFQ_REQUIRE(n >= 0 && n < (1LL << 31), "afq_rng_normal: n=%lld out of range", n);
    if (n == 0) return 0;
    AFQ_REQUIRE(key && out, "afq_rng_normal: null pointer");
    hipLaunchKernelGGL(normal_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                       (const uint32_t*)key, n, out);
    return afq::check_launch("afq_rng_normal");
}

extern "C" int afq_rng_uniform(const unsigned* key, double* out, void* stream) {
    AFQ_REQUIRE(key && out, "afq_rng_uniform: null pointer");
    hipLaunchKernelGGL(uniform_kernel, dim3(1), dim3(1), 0, (hipStream_t)stream, (const uint32_t*)key, out);
    return afq::check_launch("afq_rng_uniform");
}

extern "C" int afq_threefry2x32(unsigned k0, unsigned k1, long long n, const unsigned* c0,
                                const unsigned* c1, unsigned* o0, unsigned* o1, void* stream) {
    AFQ_REQUIRE(n >= 0, "afq_threefry2x32: n < 0");
    if (n == 0) return 0;
    AFQ_REQUIRE(c0 && c1 && o0 && o1, "afq_threefry2x32: null pointer");
    hipLaunchKernelGGL(threefry_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, k0, k1, n,
                       (const uint32_t*)c0, (const uint32_t*)c1, (uint32_t*)o0, (uint32_t*)o1);
    return afq::check_launch("afq_threefry2x32");
}

extern "C" int afq_rng_fields_grouped(int count, const afq_system_step* sys, int total_walkers,
                                      void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_rng_fields_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys, "afq_rng_fields_grouped: null pointer");
    hipLaunchKernelGGL(fields_grouped_kernel, dim3((unsigned)total_walkers), dim3(256), 0,
                       (hipStream_t)stream, count, sys);
    return afq::check_launch("afq_rng_fields_grouped");
}
