// Device building blocks shared by the single-system and the grouped kernels
// of the block end and of the device random stream: workgroup sums in a fixed
// order, the Threefry-2x32 block function with jax's key split and uniform,
// the clipped block energy and the comb.  Both forms run the same code, so
// their results agree bit for bit.
#pragma once

#include <cstdint>

#include "common.h"

namespace {

// v[k] summed over a workgroup of T threads (waves in order); the result is
// valid in thread 0.  sh: 4 * (T / 64) doubles.
template <int T>
__device__ __forceinline__ void block_sum(double (&v)[4], double* sh) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = afq::wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) sh[wave * 4 + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double s = 0.0;
            for (int q = 0; q < T / 64; ++q) s += sh[q * 4 + k];
            v[k] = s;
        }
    }
}

__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// Threefry-2x32 with 20 rounds (Salmon, Moraes, Dror & Shaw, SC 2011)
__device__ __forceinline__ void threefry2x32(uint32_t k0, uint32_t k1, uint32_t& x0, uint32_t& x1) {
    const uint32_t k2 = k0 ^ k1 ^ 0x1BD11BDAu;
    const uint32_t ks[3] = {k0, k1, k2};
    constexpr int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
    x0 += k0;
    x1 += k1;
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl32(x1, R[r & 7]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            const int s = (r + 1) >> 2;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (uint32_t)s;
        }
    }
}

// (new key, subkey) = split(key)
__device__ __forceinline__ void split_key(const uint32_t* key, uint32_t (&nk)[2], uint32_t (&sk)[2]) {
    uint32_t a0 = 0, b0 = 2, a1 = 1, b1 = 3;
    threefry2x32(key[0], key[1], a0, b0);
    threefry2x32(key[0], key[1], a1, b1);
    nk[0] = a0;
    nk[1] = a1;
    sk[0] = b0;
    sk[1] = b1;
}

// [0, 1) from 64 random bits: the top 52 bits as the mantissa of [1, 2), minus 1
__device__ __forceinline__ double unit_from_bits(uint32_t hi, uint32_t lo) {
    const uint64_t bits = ((uint64_t)hi << 32) | lo;
    return __longlong_as_double((long long)((bits >> 12) | 0x3FF0000000000000ull)) - 1.0;
}

// uniform(key) in [0, 1): one value from threefry(key, (0, 1))
__device__ __forceinline__ double uniform_value(uint32_t k0, uint32_t k1) {
    uint32_t a = 0, b = 1;
    threefry2x32(k0, k1, a, b);
    return fmax(0.0, unit_from_bits(a, b));
}

// Block energy of one walker batch (prop/blocks.py:59-79), by a workgroup of
// 1024 threads: local energies farther than thresh from e_ref count as e_ref;
// out[0] = sum w E / sum w (e_ref when sum w == 0), out[1] = sum w,
// out[2] = (1 - alpha) pop + alpha out[0].  Thread 0 reads *pop before it
// writes out, so out[2] may alias pop.  sh: 64 doubles.
__device__ __forceinline__ void block_energy_body(int nw, const double2* __restrict__ e, const double* w,
                                                  double e_ref, double thresh, double alpha, const double* pop,
                                                  double* out0, double* out1, double* out2, double* sh) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < nw; k += 1024) {
        double er = e[k].x;
        if (fabs(er - e_ref) > thresh) er = e_ref;
        acc[0] += w[k] * er;
        acc[1] += w[k];
    }
    block_sum<1024>(acc, sh);
    if (threadIdx.x == 0) {
        const double wsum = acc[1];
        const double eb = (wsum == 0.0) ? e_ref : acc[0] / wsum;
        const double p = *pop;
        *out0 = eb;
        *out1 = wsum;
        *out2 = (1.0 - alpha) * p + alpha * eb;
    }
}

// Comb indices (walkers.py:233-265) by one workgroup: the cumulative sum of
// |w| is one sequential left-to-right fp64 pass (numpy.cumsum order) into cw,
// and z_k = (tot * (k + zeta)) / nw is evaluated without contraction, so the
// indices are reproducible bit for bit.  The search uses searchsorted's order,
// in which NaN sorts above every number: with a NaN weight the total and every
// tooth are NaN, and each tooth lands on the first NaN of the cumulative sum.
// An index past the end is clamped to nw - 1, as an out-of-range gather index
// is.  wnew may alias weights (every weight is read before the first write).
// tot_s: one shared double.
__device__ __forceinline__ void comb_body(int nw, const double* weights, double zeta, double* __restrict__ cw,
                                          int* __restrict__ idx, double* wnew, double* tot_s) {
#pragma clang fp contract(off)
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) {
            s = s + fabs(weights[i]);
            cw[i] = s;
        }
        *tot_s = s;
        __threadfence();
    }
    __syncthreads();
    const double tot = *tot_s;
    const double avg = tot / (double)nw;
    for (int k = threadIdx.x; k < nw; k += blockDim.x) {
        const double kz = (double)k + zeta;
        const double prod = tot * kz;
        const double z = prod / (double)nw;
        int lo = 0, hi = nw;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const double c = cw[mid];
            if (c < z || (z != z && c == c))
                lo = mid + 1;
            else
                hi = mid;
        }
        idx[k] = lo < nw ? lo : nw - 1;
        wnew[k] = avg;
    }
}

}  // namespace
