// Block end of every system of a batch in two launches: the block energy with
// the population-shift EMA, the key split, the uniform and the comb indices
// (prop/blocks.py:59-92, walkers.py:233-265 per system) in one workgroup per
// system, then the gather of the surviving walkers, their Green's-function
// rows and log-overlaps in one workgroup per walker.  The arithmetic is that
// of the single-system kernels (device_util.h), so a batch's block end agrees
// with one system's bit for bit.
#include <cmath>
#include <cstdint>

#include "afqmc_hip.h"
#include "common.h"
#include "device_util.h"

namespace {

// Workgroup i: system i of the table.  Reads energy, weights_in, pop_in,
// key_in; writes stats[2 i], stats[2 i + 1] (block energy, weight sum),
// pop_in, key_in and weights_in in place, and the comb indices of the system
// to idx[walker0 ..] (cw[walker0 ..]: the cumulative weights).
__global__ __launch_bounds__(1024) void block_end_grouped_kernel(const afq_system_step* __restrict__ S,
                                                                 double thresh, double alpha,
                                                                 double* __restrict__ stats,
                                                                 double* __restrict__ cw,
                                                                 int* __restrict__ idx) {
    __shared__ double sh[64];
    __shared__ double tot_s;
    __shared__ double zeta_s;
    const afq_system_step& d = S[blockIdx.x];
    double* pop = const_cast<double*>(d.pop_in);
    double* w = const_cast<double*>(d.weights_in);
    // every thread has read its weights once block_sum's barrier has passed;
    // the comb writes them only after its own barrier
    block_energy_body(d.nw, (const double2*)d.energy, w, d.e_estimate, thresh, alpha, pop,
                      stats + 2 * blockIdx.x, stats + 2 * blockIdx.x + 1, pop, sh);
    if (threadIdx.x == 0) {
        uint32_t* key = const_cast<uint32_t*>((const uint32_t*)d.key_in);
        uint32_t nk[2], sk[2];
        split_key(key, nk, sk);
        key[0] = nk[0];
        key[1] = nk[1];
        zeta_s = uniform_value(sk[0], sk[1]);
    }
    __syncthreads();
    comb_body(d.nw, w, zeta_s, cw + d.walker0, idx + d.walker0, w, &tot_s);
}

// Workgroup b: walker row k = b - walker0 of its system, j = idx[walker0 + k]:
// walkers_in[k] = walkers_out[j], both planes of the Green's-function rows
// (greens_dst + (greens - greens_src))[p, k] = greens[p, j], and
// log_overlaps_in[k] = 2 logdet[j].
__global__ __launch_bounds__(256) void gather_grouped_kernel(int count, const afq_system_step* __restrict__ S,
                                                             const int* __restrict__ idx,
                                                             const double* greens_src,
                                                             double* __restrict__ greens_dst) {
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    const int k = b - d.walker0;
    const long long j = idx[b];
    const long long wl = (long long)d.n * d.o;  // complex entries per walker
    const double2* src = (const double2*)d.walkers_out + j * wl;
    double2* dst = (double2*)const_cast<double*>(d.walkers_in) + (long long)k * wl;
    for (long long e = threadIdx.x; e < wl; e += 256) dst[e] = src[e];
    const long long gl = wl;                    // o * n per plane row
    const long long plane = (long long)d.nw * gl;
    const double* gs = d.greens;
    double* gd = greens_dst + (d.greens - greens_src);
    for (int p = 0; p < 2; ++p)
        for (long long e = threadIdx.x; e < gl; e += 256) gd[p * plane + k * gl + e] = gs[p * plane + j * gl + e];
    if (threadIdx.x == 0) {
        const double2 ld = ((const double2*)d.logdet)[j];
        ((double2*)const_cast<double*>(d.log_overlaps_in))[k] = make_double2(2.0 * ld.x, 2.0 * ld.y);
    }
}

}  // namespace

extern "C" int afq_block_end_grouped(int count, const afq_system_step* sys, double thresh, double shift_ema,
                                     double* stats, double* cw, int* idx, void* stream) {
    AFQ_REQUIRE(count > 0, "afq_block_end_grouped: count=%d", count);
    AFQ_REQUIRE(sys && stats && cw && idx, "afq_block_end_grouped: null pointer");
    hipLaunchKernelGGL(block_end_grouped_kernel, dim3((unsigned)count), dim3(1024), 0, (hipStream_t)stream, sys,
                       thresh, shift_ema, stats, cw, idx);
    return afq::check_launch("afq_block_end_grouped");
}

extern "C" int afq_reconfigure_gather_grouped(int count, const afq_system_step* sys, int total_walkers,
                                              const int* idx, const double* greens_src, double* greens_dst,
                                              void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_reconfigure_gather_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys && idx && greens_src && greens_dst && greens_src != greens_dst,
                "afq_reconfigure_gather_grouped: null or aliased pointer");
    hipLaunchKernelGGL(gather_grouped_kernel, dim3((unsigned)total_walkers), dim3(256), 0, (hipStream_t)stream,
                       count, sys, idx, greens_src, greens_dst);
    return afq::check_launch("afq_reconfigure_gather_grouped");
}
