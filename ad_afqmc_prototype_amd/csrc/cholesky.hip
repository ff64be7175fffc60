// Modified (pivoted, incomplete) Cholesky decomposition of the two-electron
// integral matrix into Cholesky vectors: Hamiltonian ingestion
// (integrals.py:3-46).  mat is the (size x size) matrix of the 4-fold
// symmetric ERIs in packed pair indices (size = n (n + 1) / 2, pair (m, p),
// m >= p, at m (m + 1) / 2 + p).
//
// The iteration is the reference's:
//   start:  nu = argmax diag, d = diag[nu], v_0 = mat[nu] / sqrt(d), M = 0;
//   while |d| > max_error and nchol + 1 < max_vectors:
//       M += v_nchol^2;  nu = argmax |diag - M|;  d = |diag - M|[nu];
//       v_{nchol+1} = (mat[nu] - sum_{i <= nchol} v_i[nu] v_i) / sqrt(d + 1e-10);
//       nchol += 1
//   result: v_0 .. v_{nchol-1}.
// Every iteration is three launches that read the iteration index, pivot and
// stop flag from a small device state, so a caller can enqueue many
// iterations without synchronising and look at the stop flag now and then;
// iterations enqueued after the stop do nothing.  argmax takes the first
// index among equal values (numpy's rule).  The vectors' k-sums run in a fixed
// sequential order.
#include <cmath>

#include "afqmc_hip.h"
#include "common.h"

namespace {

constexpr int THREADS = 256;
constexpr int MAX_PARTIAL = 1024;

// state: [0] d (pivot value), [1] nu, [2] nchol, [3] done
__device__ __forceinline__ void arg_better(double v, int i, double& bv, int& bi) {
    if (v > bv || (v == bv && i < bi)) {
        bv = v;
        bi = i;
    }
}

// Workgroup-level (max, first index) of val(j) over j, written to partial[blockIdx].
template <typename F>
__device__ __forceinline__ void block_argmax(int size, F val, double* partial) {
    __shared__ double sv[THREADS / 64];
    __shared__ int si[THREADS / 64];
    // start at index 0 so that the pivot is a valid index even when no value
    // compares greater than -inf (all NaN): the iteration then yields NaN
    // vectors and stops at max_vectors, but reads only in bounds
    double bv = -INFINITY;
    int bi = 0;
    for (int j = blockIdx.x * THREADS + threadIdx.x; j < size; j += gridDim.x * THREADS) arg_better(val(j), j, bv, bi);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        arg_better(ov, oi, bv, bi);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        sv[wave] = bv;
        si[wave] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < THREADS / 64; ++w) arg_better(sv[w], si[w], bv, bi);
        partial[2 * blockIdx.x] = bv;
        partial[2 * blockIdx.x + 1] = (double)bi;
    }
}

// start: argmax of the (signed) diagonal
__global__ __launch_bounds__(THREADS) void mcd_init_partial(int size, const double* __restrict__ mat, long long ldm,
                                                            double* __restrict__ resid, double* __restrict__ partial) {
    for (int j = blockIdx.x * THREADS + threadIdx.x; j < size; j += gridDim.x * THREADS) resid[j] = 0.0;
    block_argmax(size, [&](int j) { return mat[(long long)j * ldm + j]; }, partial);
}

// one iteration, part 1: M += v_k^2 and argmax |diag - M|
__global__ __launch_bounds__(THREADS) void mcd_pivot_partial(int size, const double* __restrict__ mat, long long ldm,
                                                             const double* __restrict__ vecs,
                                                             double* __restrict__ resid, double* __restrict__ partial,
                                                             const double* __restrict__ state) {
    if (state[3] != 0.0) return;
    const long long k = (long long)state[2];
    const double* vk = vecs + k * size;
    block_argmax(size, [&](int j) {
        const double m = resid[j] + vk[j] * vk[j];
        resid[j] = m;
        return fabs(mat[(long long)j * ldm + j] - m);
    }, partial);
}

// reduce the partials; at the start (init != 0) also decide whether to iterate
__global__ __launch_bounds__(64) void mcd_pivot_final(int nparts, const double* __restrict__ partial,
                                                      double* __restrict__ state, double max_error,
                                                      int max_vectors, int init) {
    if (!init && state[3] != 0.0) return;
    double bv = -INFINITY;
    int bi = 0;
    for (int p = threadIdx.x; p < nparts; p += 64) arg_better(partial[2 * p], (int)partial[2 * p + 1], bv, bi);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(bv, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        arg_better(ov, oi, bv, bi);
    }
    if (threadIdx.x == 0) {
        if (init) {
            state[0] = bv;
            state[1] = (double)bi;
            state[2] = 0.0;
            // the loop test before the first iteration
            state[3] = (fabs(bv) > max_error && 1 < max_vectors) ? 0.0 : 1.0;
        } else {
            const double nchol = state[2] + 1.0;  // v_{k+1} is about to be formed
            state[0] = bv;
            state[1] = (double)bi;
            state[2] = nchol;
            // the loop test of the next iteration; a vector formed after a
            // failed test is not part of the result, so it is not formed
            state[3] = (fabs(bv) > max_error && nchol + 1.0 < (double)max_vectors) ? 0.0 : 1.0;
        }
    }
}

// v_0 = mat[nu] / sqrt(d)
__global__ __launch_bounds__(THREADS) void mcd_first_vector(int size, const double* __restrict__ mat, long long ldm,
                                                            double* __restrict__ vecs, const double* __restrict__ state) {
    const long long nu = (long long)state[1];
    const double s = sqrt(state[0]);
    for (int j = blockIdx.x * THREADS + threadIdx.x; j < size; j += gridDim.x * THREADS)
        vecs[j] = mat[nu * ldm + j] / s;
}

// one iteration, part 2: v_{k+1} = (mat[nu] - sum_{i <= k} v_i[nu] v_i) / sqrt(d + 1e-10)
__global__ __launch_bounds__(THREADS) void mcd_new_vector(int size, const double* __restrict__ mat, long long ldm,
                                                          double* __restrict__ vecs, const double* __restrict__ state) {
    if (state[3] != 0.0) return;
    const long long k1 = (long long)state[2];  // index of the new vector
    const long long nu = (long long)state[1];
    const double s = sqrt(state[0] + 1.0e-10);
    for (int j = blockIdx.x * THREADS + threadIdx.x; j < size; j += gridDim.x * THREADS) {
        double r = 0.0;
        for (long long i = 0; i < k1; ++i) r += vecs[i * size + nu] * vecs[i * size + j];
        vecs[k1 * size + j] = (mat[nu * ldm + j] - r) / s;
    }
}

__global__ __launch_bounds__(THREADS) void unpack_pairs_kernel(int nchol, int n, const double* __restrict__ vecs,
                                                               long long ldv, double* __restrict__ chol) {
    const long long total = (long long)nchol * n * n;
    for (long long e = (long long)blockIdx.x * THREADS + threadIdx.x; e < total;
         e += (long long)gridDim.x * THREADS) {
        const long long i = e / ((long long)n * n);
        const int r = (int)(e - i * n * n);
        const int m = r / n, p = r - (r / n) * n;
        const int a = m > p ? m : p, b = m > p ? p : m;
        chol[e] = vecs[i * ldv + (long long)a * (a + 1) / 2 + b];
    }
}

int parts_for(int size) {
    const int b = afq::ceil_div(size, THREADS);
    return b > MAX_PARTIAL ? MAX_PARTIAL : b;
}

}  // namespace

extern "C" int afq_mcd_workspace(int size, long long* resid_doubles, long long* partial_doubles) {
    AFQ_REQUIRE(size > 0 && resid_doubles && partial_doubles, "afq_mcd_workspace: bad arguments");
    *resid_doubles = size;
    *partial_doubles = 2LL * parts_for(size);
    return 0;
}

extern "C" int afq_mcd_start(int size, const double* mat, long long ldm, double max_error, int max_vectors,
                             double* vecs, double* resid, double* partial, double* state, void* stream) {
    AFQ_REQUIRE(size > 0 && ldm >= size, "afq_mcd_start: size=%d ldm=%lld", size, ldm);
    AFQ_REQUIRE(max_vectors >= 1 && max_vectors <= size, "afq_mcd_start: max_vectors=%d not in 1..%d",
                max_vectors, size);
    AFQ_REQUIRE(mat && vecs && resid && partial && state, "afq_mcd_start: null pointer");
    hipStream_t st = (hipStream_t)stream;
    const int np = parts_for(size);
    hipLaunchKernelGGL(mcd_init_partial, dim3(np), dim3(THREADS), 0, st, size, mat, ldm, resid, partial);
    hipLaunchKernelGGL(mcd_pivot_final, dim3(1), dim3(64), 0, st, np, partial, state, max_error, max_vectors, 1);
    hipLaunchKernelGGL(mcd_first_vector, dim3(np), dim3(THREADS), 0, st, size, mat, ldm, vecs, state);
    return afq::check_launch("afq_mcd_start");
}

extern "C" int afq_mcd_iterate(int size, const double* mat, long long ldm, double max_error, int max_vectors,
                               int iterations, double* vecs, double* resid, double* partial, double* state,
                               void* stream) {
    AFQ_REQUIRE(size > 0 && ldm >= size, "afq_mcd_iterate: size=%d ldm=%lld", size, ldm);
    AFQ_REQUIRE(max_vectors >= 1 && max_vectors <= size, "afq_mcd_iterate: max_vectors=%d not in 1..%d",
                max_vectors, size);
    AFQ_REQUIRE(iterations >= 0, "afq_mcd_iterate: iterations < 0");
    AFQ_REQUIRE(mat && vecs && resid && partial && state, "afq_mcd_iterate: null pointer");
    hipStream_t st = (hipStream_t)stream;
    const int np = parts_for(size);
    for (int it = 0; it < iterations; ++it) {
        hipLaunchKernelGGL(mcd_pivot_partial, dim3(np), dim3(THREADS), 0, st, size, mat, ldm, vecs, resid, partial,
                           state);
        hipLaunchKernelGGL(mcd_pivot_final, dim3(1), dim3(64), 0, st, np, partial, state, max_error, max_vectors, 0);
        hipLaunchKernelGGL(mcd_new_vector, dim3(np), dim3(THREADS), 0, st, size, mat, ldm, vecs, state);
    }
    return afq::check_launch("afq_mcd_iterate");
}

extern "C" int afq_unpack_pairs(int nchol, int n, const double* vecs, long long ldv, double* chol, void* stream) {
    AFQ_REQUIRE(nchol >= 0 && n > 0 && ldv >= (long long)n * (n + 1) / 2, "afq_unpack_pairs: bad sizes");
    if (nchol == 0) return 0;
    AFQ_REQUIRE(vecs && chol, "afq_unpack_pairs: null pointer");
    const long long total = (long long)nchol * n * n;
    long long blocks = (total + THREADS - 1) / THREADS;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(unpack_pairs_kernel, dim3((unsigned)blocks), dim3(THREADS), 0, (hipStream_t)stream, nchol,
                       n, vecs, ldv, chol);
    return afq::check_launch("afq_unpack_pairs");
}
