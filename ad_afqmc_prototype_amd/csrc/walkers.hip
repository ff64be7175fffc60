// Walker-container kernels: orthonormalization (QR), comb reconfiguration
// indices and the row gather.
#include <cmath>

#include <mutex>

#include "afqmc_hip.h"
#include "common.h"
#include "device_util.h"

namespace {

constexpr int THREADS = 256;

// ---------------------------------------------------------------------------
// Reduced QR by classical Gram-Schmidt with one re-orthogonalisation pass
// (CGS2).  R_jj = ||v_j|| is real and positive, so Q is the unique Q of the
// diag(R) >= 0 convention of walkers.py:169-180 (LAPACK geqrf + phase fix),
// equal to it to rounding for numerically full-rank walkers.  One workgroup per
// walker; the walker (n x o complex) sits in LDS with an odd row pitch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void qr_body(int n, int o, const double2* __restrict__ W,
                                        double2* __restrict__ Q, double2* q) {
    __shared__ double2 r[64];
    __shared__ double scale_s;
    const int pitch = o + 1;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    for (int e = tid; e < n * o; e += THREADS) {
        const int p = e / o;
        const int j = e - p * o;
        q[p * pitch + j] = W[e];
    }
    __syncthreads();

    for (int j = 0; j < o; ++j) {
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = wave; i < j; i += 4) {
                double sr = 0.0, si = 0.0;
                for (int p = lane; p < n; p += 64) {
                    const double2 a = q[p * pitch + i];
                    const double2 b = q[p * pitch + j];
                    sr += a.x * b.x + a.y * b.y;  // conj(a) b
                    si += a.x * b.y - a.y * b.x;
                }
                sr = afq::wave_sum(sr);
                si = afq::wave_sum(si);
                if (lane == 0) r[i] = make_double2(sr, si);
            }
            __syncthreads();
            for (int p = tid; p < n; p += THREADS) {
                double2 v = q[p * pitch + j];
                for (int i = 0; i < j; ++i) {
                    const double2 a = q[p * pitch + i];
                    const double2 c = r[i];
                    v.x -= a.x * c.x - a.y * c.y;
                    v.y -= a.x * c.y + a.y * c.x;
                }
                q[p * pitch + j] = v;
            }
            __syncthreads();
        }
        if (wave == 0) {
            double s = 0.0;
            for (int p = lane; p < n; p += 64) {
                const double2 b = q[p * pitch + j];
                s += b.x * b.x + b.y * b.y;
            }
            s = afq::wave_sum(s);
            if (lane == 0) scale_s = 1.0 / sqrt(s);
        }
        __syncthreads();
        const double sc = scale_s;
        for (int p = tid; p < n; p += THREADS) {
            double2 v = q[p * pitch + j];
            v.x *= sc;
            v.y *= sc;
            q[p * pitch + j] = v;
        }
        __syncthreads();
    }

    for (int e = tid; e < n * o; e += THREADS) {
        const int p = e / o;
        const int j = e - p * o;
        Q[e] = q[p * pitch + j];
    }
}

__global__ __launch_bounds__(THREADS) void qr_kernel(int n, int o, const double2* __restrict__ win,
                                                      double2* __restrict__ wout) {
    extern __shared__ double2 q_lds[];
    const long long off = (long long)blockIdx.x * n * o;
    qr_body(n, o, win + off, wout + off, q_lds);
}

// One workgroup per walker of every system: walkers_in -> walkers_out.
__global__ __launch_bounds__(THREADS) void qr_grouped_kernel(int count,
                                                             const afq_system_step* __restrict__ S) {
    extern __shared__ double2 q_lds[];
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    const long long off = (long long)(b - d.walker0) * d.n * d.o;
    qr_body(d.n, d.o, (const double2*)d.walkers_in + off, (double2*)d.walkers_out + off, q_lds);
}

// ---------------------------------------------------------------------------
// Reduced QR by two Cholesky-QR passes (CholQR2) for o <= 32, one wave per
// walker, four per workgroup: W^H W = R^H R (Gram matrix on fp64 MFMA,
// Cholesky in the wave's LDS, R upper triangular with a real positive
// diagonal), Q = W R^-1 row by row; the second pass on Q restores
// orthogonality to rounding.  The Q of the diag(R) > 0 convention is unique,
// so this is the Q of walkers.py:169-180 to rounding for walkers far from
// rank deficiency (condition number << 1e8; walkers are re-orthonormalised
// every few steps).  One pass per launch (the second in place: each lane reads
// its rows before writing them, and the Gram matrix is formed before any row
// is written).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

template <int MAXO>
struct cholqr_lds {
    double2 g[MAXO][MAXO + 1];
    double rinv[MAXO];
};

template <int MAXO>
__device__ __forceinline__ void cholqr_pass(int n, int o, const double2* W, double2* Q, cholqr_lds<MAXO>& S) {
    constexpr int NT = (MAXO + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    // 1. Gram matrix g = W^H W
    {
        const int nkb = (n + 15) >> 4;
        f64x4 acc[NT][NT][2];
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) acc[ti][tj][0] = acc[ti][tj][1] = f64x4{0.0, 0.0, 0.0, 0.0};
        double2 x[NT][4], xn[NT][4];
        auto load = [&](int kb, double2 (&xv)[NT][4]) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int p = kb * 16 + lk + kk;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int c = t * 16 + lr;
                    xv[t][kk] = (p < n && c < o) ? W[p * o + c] : make_double2(0.0, 0.0);
                }
            }
        };
        load(0, x);
        for (int kb = 0; kb < nkb; ++kb) {
            load(kb + 1, xn);
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        const double2 a = x[ti][kk], b = x[tj][kk];
                        acc[ti][tj][0] = afq::mfma_f64(a.x, b.x, acc[ti][tj][0]);
                        acc[ti][tj][0] = afq::mfma_f64(a.y, b.y, acc[ti][tj][0]);
                        acc[ti][tj][1] = afq::mfma_f64(a.x, b.y, acc[ti][tj][1]);
                        acc[ti][tj][1] = afq::mfma_f64(-a.y, b.x, acc[ti][tj][1]);
                    }
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) x[t][kk] = xn[t][kk];
        }
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = ti * 16 + afq::mfma16<double>::row(lane, r);
                    const int j = tj * 16 + lr;
                    if (i < o && j < o) S.g[i][j] = make_double2(acc[ti][tj][0][r], acc[ti][tj][1][r]);
                }
    }
    wave_lds_sync();
    // 2. Cholesky g = R^H R, R over the upper triangle (right-looking)
    for (int k = 0; k < o; ++k) {
        const double d = sqrt(S.g[k][k].x);
        const double dinv = 1.0 / d;
        for (int j = k + lane; j < o; j += 64) {
            const double2 v = S.g[k][j];
            S.g[k][j] = (j == k) ? make_double2(d, 0.0) : make_double2(v.x * dinv, v.y * dinv);
        }
        if (lane == 0) S.rinv[k] = dinv;
        wave_lds_sync();
        // trailing update of the upper triangle: lane & 31 is the column j,
        // lane >> 5 the row parity (o <= 32: no index division)
        {
            const int j = k + 1 + (lane & 31);
            if (j < o) {
                const double2 b = S.g[k][j];
                for (int i = k + 1 + (lane >> 5); i <= j; i += 2) {
                    const double2 a = S.g[k][i];  // g_ij -= conj(a) b, as four FMAs
                    double2 g = S.g[i][j];
                    g.x = fma(-a.x, b.x, fma(-a.y, b.y, g.x));
                    g.y = fma(-a.x, b.y, fma(a.y, b.x, g.y));
                    S.g[i][j] = g;
                }
            }
        }
        wave_lds_sync();
    }
    // 3. rows: x R = w_p
    for (int p = lane; p < n; p += 64) {
        asm volatile("" ::: "memory");
        double2 xr[MAXO];
#pragma unroll
        for (int j = 0; j < MAXO; ++j) xr[j] = (j < o) ? W[p * o + j] : make_double2(0.0, 0.0);
#pragma unroll
        for (int i = 0; i < MAXO; ++i) {
            if (i < o) {
                const double ri = S.rinv[i];
                xr[i].x *= ri;
                xr[i].y *= ri;
#pragma unroll
                for (int j = i + 1; j < MAXO; ++j) {
                    if (j < o) {
                        // x_j -= x_i r_ij, as four FMAs (no separately rounded product)
                        const double2 r = S.g[i][j];
                        xr[j].x = fma(-xr[i].x, r.x, fma(xr[i].y, r.y, xr[j].x));
                        xr[j].y = fma(-xr[i].x, r.y, fma(-xr[i].y, r.x, xr[j].y));
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < MAXO; ++j)
            if (j < o) Q[p * o + j] = xr[j];
    }
}

template <int MAXO>
__global__ __launch_bounds__(THREADS) void cholqr_kernel(int nw, int n, int o, const double2* src,
                                                         double2* dst) {  // src may equal dst
    __shared__ cholqr_lds<MAXO> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = blockIdx.x * (THREADS / 64) + wave;
    if (w >= nw) return;
    const long long off = (long long)w * n * o;
    cholqr_pass<MAXO>(n, o, src + off, dst + off, L[wave]);
}

// second == 0: walkers_in -> walkers_out; second == 1: walkers_out in place
template <int MAXO>
__global__ __launch_bounds__(THREADS) void cholqr_grouped_kernel(int count, int total, int second,
                                                                 const afq_system_step* __restrict__ S) {
    __shared__ cholqr_lds<MAXO> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = blockIdx.x * (THREADS / 64) + wave;
    if (b >= total) return;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    const long long off = (long long)(b - d.walker0) * d.n * d.o;
    const double2* src = (second ? (const double2*)d.walkers_out : (const double2*)d.walkers_in) + off;
    cholqr_pass<MAXO>(d.n, d.o, src, (double2*)d.walkers_out + off, L[wave]);
}

template <int MAXO>
void launch_cholqr(int nw, int n, int o, const double2* win, double2* wout, hipStream_t st) {
    const dim3 grid((unsigned)afq::ceil_div(nw, THREADS / 64));
    hipLaunchKernelGGL(cholqr_kernel<MAXO>, grid, dim3(THREADS), 0, st, nw, n, o, win, wout);
    hipLaunchKernelGGL(cholqr_kernel<MAXO>, grid, dim3(THREADS), 0, st, nw, n, o, (const double2*)wout, wout);
}

template <int MAXO>
void launch_cholqr_grouped(int count, int total, const afq_system_step* sys, hipStream_t st) {
    const dim3 grid((unsigned)afq::ceil_div(total, THREADS / 64));
    hipLaunchKernelGGL(cholqr_grouped_kernel<MAXO>, grid, dim3(THREADS), 0, st, count, total, 0, sys);
    hipLaunchKernelGGL(cholqr_grouped_kernel<MAXO>, grid, dim3(THREADS), 0, st, count, total, 1, sys);
}

// Comb indices of one walker batch (comb_body): zeta from the argument, or
// from the device when zeta_p is set.
__global__ __launch_bounds__(1024) void comb_kernel(int nw, const double* __restrict__ weights,
                                                     double zeta_v, const double* __restrict__ zeta_p,
                                                     double* __restrict__ cw,
                                                     int* __restrict__ idx,
                                                     double* __restrict__ wnew) {
    __shared__ double tot_s;
    comb_body(nw, weights, zeta_p ? zeta_p[0] : zeta_v, cw, idx, wnew, &tot_s);
}

__global__ void gather_kernel(int nrows, long long row_len, int nplanes, long long sps,
                              long long dps, const int* __restrict__ idx,
                              const double* __restrict__ src, double* __restrict__ dst) {
    const long long per_plane = (long long)nrows * row_len;
    const long long total = per_plane * nplanes;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long p = e / per_plane;
        const long long rem = e - p * per_plane;
        const long long k = rem / row_len;
        const long long c = rem - k * row_len;
        dst[p * dps + k * row_len + c] = src[p * sps + (long long)idx[k] * row_len + c];
    }
}

}  // namespace

extern "C" int afq_orthonormalize(int nw, int n, int o, const double* win, double* wout,
                                  void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_orthonormalize: bad sizes nw=%d n=%d o=%d", nw, n, o);
    AFQ_REQUIRE(o <= 64, "afq_orthonormalize: o=%d exceeds the supported maximum of 64", o);
    AFQ_REQUIRE(win && wout && win != wout, "afq_orthonormalize: null or aliased pointer");
    hipStream_t st = (hipStream_t)stream;
    if (o <= 32) {
        if (o <= 8)
            launch_cholqr<8>(nw, n, o, (const double2*)win, (double2*)wout, st);
        else if (o <= 16)
            launch_cholqr<16>(nw, n, o, (const double2*)win, (double2*)wout, st);
        else
            launch_cholqr<32>(nw, n, o, (const double2*)win, (double2*)wout, st);
        return afq::check_launch("afq_orthonormalize");
    }
    const size_t lds = (size_t)n * (o + 1) * sizeof(double2);
    AFQ_REQUIRE(lds + 64 * sizeof(double2) + 64 <= 160 * 1024,
                "afq_orthonormalize: walker n=%d o=%d does not fit in LDS", n, o);
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)qr_kernel);
    hipLaunchKernelGGL(qr_kernel, dim3((unsigned)nw), dim3(THREADS), lds, (hipStream_t)stream, n, o,
                       (const double2*)win, (double2*)wout);
    return afq::check_launch("afq_orthonormalize");
}

extern "C" int afq_comb_indices(int nw, const double* weights, double zeta, double* cumsum,
                                int* idx, double* weights_new, void* stream) {
    AFQ_REQUIRE(nw > 0, "afq_comb_indices: nw=%d", nw);
    AFQ_REQUIRE(weights && cumsum && idx && weights_new, "afq_comb_indices: null pointer");
    hipLaunchKernelGGL(comb_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, nw, weights, zeta,
                       (const double*)nullptr, cumsum, idx, weights_new);
    return afq::check_launch("afq_comb_indices");
}

extern "C" int afq_comb_indices_dev(int nw, const double* weights, const double* zeta,
                                    double* cumsum, int* idx, double* weights_new, void* stream) {
    AFQ_REQUIRE(nw > 0, "afq_comb_indices_dev: nw=%d", nw);
    AFQ_REQUIRE(weights && zeta && cumsum && idx && weights_new, "afq_comb_indices_dev: null pointer");
    hipLaunchKernelGGL(comb_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, nw, weights, 0.0, zeta,
                       cumsum, idx, weights_new);
    return afq::check_launch("afq_comb_indices_dev");
}

// ---------------------------------------------------------------------------
// Comb indices without caller scratch: the cumulative sum lives in a
// file-scope device array (one copy per GPU, held until the process ends).
// Its address is looked up once per device with that device current; the
// kernel receives that address, never the array's host-side name.  Every call
// on one GPU shares the array: calls that overlap on one GPU (different
// streams, or graphs replayed concurrently) race on it.  Nothing in it has to
// start at zero: the kernel writes cw[0..nw) before any lane reads it.
// ---------------------------------------------------------------------------
namespace {
constexpr int kCombScratchWalkers = 1 << 18;  // 2 MiB of fp64
constexpr int kMaxDevices = 64;
__device__ __align__(256) double scratch[kCombScratchWalkers];
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int comb_scratch(hipStream_t st, double** out) {
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) {
        afq::set_error("afq_comb_indices_noalloc: hipGetDevice failed");
        return -2;
    }
    int dev = cur;
    if (st != nullptr && hipStreamGetDevice(st, &dev) != hipSuccess) {
        afq::set_error("afq_comb_indices_noalloc: hipStreamGetDevice failed");
        return -2;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        afq::set_error("afq_comb_indices_noalloc: device %d beyond %d", dev, kMaxDevices);
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        if (dev != cur && hipSetDevice(dev) != hipSuccess) {
            afq::set_error("afq_comb_indices_noalloc: hipSetDevice(%d) failed", dev);
            return -2;
        }
        void* ptr = nullptr;
        const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (dev != cur) (void)hipSetDevice(cur);
        if (err != hipSuccess || ptr == nullptr) {
            afq::set_error("afq_comb_indices_noalloc: hipGetSymbolAddress failed: %s",
                           hipGetErrorString(err));
            return -2;
        }
        g_scratch_ptr[dev] = (double*)ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" int afq_comb_indices_noalloc(int nw, const double* weights, const double* zeta, int* idx,
                                        double* weights_new, void* stream) {
    AFQ_REQUIRE(nw > 0, "afq_comb_indices_noalloc: nw=%d", nw);
    AFQ_REQUIRE(nw <= kCombScratchWalkers,
                "afq_comb_indices_noalloc: nw=%d exceeds the library's scratch of %d walkers", nw,
                kCombScratchWalkers);
    AFQ_REQUIRE(weights && zeta && idx && weights_new, "afq_comb_indices_noalloc: null pointer");
    double* cw = nullptr;
    const int rc = comb_scratch((hipStream_t)stream, &cw);
    if (rc != 0) return rc;
    hipLaunchKernelGGL(comb_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, nw, weights, 0.0, zeta,
                       cw, idx, weights_new);
    return afq::check_launch("afq_comb_indices_noalloc");
}

extern "C" int afq_gather_rows(int nrows, long long row_len, int nplanes, long long src_plane_stride,
                               long long dst_plane_stride, const int* idx, const double* src,
                               double* dst, void* stream) {
    AFQ_REQUIRE(nrows > 0 && row_len > 0 && nplanes > 0, "afq_gather_rows: bad sizes");
    AFQ_REQUIRE(idx && src && dst && src != dst, "afq_gather_rows: null or aliased pointer");
    const long long total = (long long)nrows * row_len * nplanes;
    long long blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       nrows, row_len, nplanes, src_plane_stride, dst_plane_stride, idx, src, dst);
    return afq::check_launch("afq_gather_rows");
}

extern "C" int afq_orthonormalize_grouped(int count, const afq_system_step* sys, int total_walkers,
                                          int max_n, int max_o, void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_orthonormalize_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys, "afq_orthonormalize_grouped: null pointer");
    AFQ_REQUIRE(max_o > 0 && max_o <= 64 && max_n >= max_o, "afq_orthonormalize_grouped: max_n=%d max_o=%d",
                max_n, max_o);
    if (max_o <= 32) {
        hipStream_t st = (hipStream_t)stream;
        if (max_o <= 8)
            launch_cholqr_grouped<8>(count, total_walkers, sys, st);
        else if (max_o <= 16)
            launch_cholqr_grouped<16>(count, total_walkers, sys, st);
        else
            launch_cholqr_grouped<32>(count, total_walkers, sys, st);
        return afq::check_launch("afq_orthonormalize_grouped");
    }
    const size_t lds = (size_t)max_n * (max_o + 1) * sizeof(double2);
    AFQ_REQUIRE(lds + 64 * sizeof(double2) + 64 <= 160 * 1024,
                "afq_orthonormalize_grouped: max_n=%d max_o=%d needs more than 160 KB of LDS",
                max_n, max_o);
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)qr_grouped_kernel);
    hipLaunchKernelGGL(qr_grouped_kernel, dim3((unsigned)total_walkers), dim3(THREADS), lds,
                       (hipStream_t)stream, count, sys);
    return afq::check_launch("afq_orthonormalize_grouped");
}
