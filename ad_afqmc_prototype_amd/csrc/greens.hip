// Overlap matrix, LU log-determinant and half-rotated Green's function.
//
// One 256-thread workgroup per (walker, spin):
//   1. m = C^T w (o x o complex) into LDS;
//   2. LU with partial pivoting on |re| + |im| (the LAPACK zgetrf pivot rule
//      used by jnp.linalg.det / jnp.linalg.solve), multipliers formed with the
//      pivot's reciprocal;
//   3. log det m = sum log(U_kk) + i pi (#row swaps);
//   4. each thread owns rows p of w and solves x m = w_p in registers:
//      u U = w_p (forward), v L = u (backward), x[perm[k]] = v[k];
//      G[i][p] = x_i is written to the split-plane greens buffer, coalesced
//      along p.
#include <cmath>

#include "afqmc_hip.h"
#include "common.h"

namespace {

constexpr int THREADS = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// y -= a b as four fused multiply-adds (the product is not rounded on its own)
__device__ __forceinline__ void csub_mul(double2& y, double2 a, double2 b) {
    y.x = fma(-a.x, b.x, fma(a.y, b.y, y.x));
    y.y = fma(-a.x, b.y, fma(-a.y, b.x, y.y));
}

__device__ __forceinline__ double2 crecip(double2 a) {
    // 1/a by the scaled formula (Smith's method) to avoid overflow.
    if (fabs(a.x) >= fabs(a.y)) {
        const double r = a.y / a.x;
        const double d = a.x + a.y * r;
        return make_double2(1.0 / d, -r / d);
    }
    const double r = a.x / a.y;
    const double d = a.x * r + a.y;
    return make_double2(r / d, -1.0 / d);
}

// One (walker, spin): W (n x o complex), Cm (n x o real) -> G^T rows into
// gre / gim (o x n each), log det into *ld.
template <int MAXO>
__device__ __forceinline__ void greens_body(int n, int o, const double2* __restrict__ W,
                                            const double* __restrict__ Cm, double* __restrict__ gre,
                                            double* __restrict__ gim, double2* __restrict__ ld) {
    __shared__ double2 m[MAXO][MAXO + 1];
    __shared__ double2 dinv[MAXO];
    __shared__ int perm[MAXO];
    __shared__ int piv_s;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;

    // 1. m = C^T w on v_mfma_f64_16x16x4: each wave owns 16x16 tiles of m;
    //    A = C^T (lane: row i, four consecutive p), B = w (same p, column j).
    {
        const int ntile = (o + 15) >> 4;
        const int lr = lane & 15;
        const int lk = (lane >> 4) * 4;
        const int nkb = (n + 15) >> 4;
        for (int t = wave; t < ntile * ntile; t += THREADS / 64) {
            const int it = t / ntile;
            const int jt = t - it * ntile;
            const int i = it * 16 + lr;
            const int j = jt * 16 + lr;
            f64x4 ar = {0.0, 0.0, 0.0, 0.0}, ai = {0.0, 0.0, 0.0, 0.0};
            for (int kb = 0; kb < nkb; ++kb) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int p = kb * 16 + lk + kk;
                    const bool okp = p < n;
                    const double a = (okp && i < o) ? Cm[p * o + i] : 0.0;
                    const double2 x = (okp && j < o) ? W[p * o + j] : make_double2(0.0, 0.0);
                    ar = afq::mfma_f64(a, x.x, ar);
                    ai = afq::mfma_f64(a, x.y, ai);
                }
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ii = it * 16 + afq::mfma16<double>::row(lane, r);
                if (ii < o && j < o) m[ii][j] = make_double2(ar[r], ai[r]);
            }
        }
    }
    if (tid < o) perm[tid] = tid;
    __syncthreads();

    // 2. LU with partial pivoting
    int nswap = 0;
    for (int k = 0; k < o; ++k) {
        if (wave == 0) {
            double best = -1.0;
            int bi = k;
            for (int r = k + lane; r < o; r += 64) {
                const double v = fabs(m[r][k].x) + fabs(m[r][k].y);
                if (v > best) {
                    best = v;
                    bi = r;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (lane == 0) piv_s = bi;
        }
        __syncthreads();
        const int piv = piv_s;
        if (piv != k) {
            for (int c = tid; c < o; c += THREADS) {
                const double2 t = m[k][c];
                m[k][c] = m[piv][c];
                m[piv][c] = t;
            }
            if (tid == 0) {
                const int t = perm[k];
                perm[k] = perm[piv];
                perm[piv] = t;
            }
            ++nswap;
        }
        __syncthreads();
        const double2 inv = crecip(m[k][k]);
        for (int r = k + 1 + tid; r < o; r += THREADS) m[r][k] = cmul(m[r][k], inv);
        __syncthreads();
        const int rem = o - k - 1;
        for (int e = tid; e < rem * rem; e += THREADS) {
            const int r = k + 1 + e / rem;
            const int c = k + 1 + e % rem;
            double2 y = m[r][c];
            csub_mul(y, m[r][k], m[k][c]);
            m[r][c] = y;
        }
        __syncthreads();
    }

    // 3. log det and pivot reciprocals
    if (tid < o) dinv[tid] = crecip(m[tid][tid]);
    if (tid == 0) {
        double lr = 0.0, li = 0.0;
        for (int k = 0; k < o; ++k) {
            lr += log(hypot(m[k][k].x, m[k][k].y));
            li += atan2(m[k][k].y, m[k][k].x);
        }
        if (nswap & 1) li += M_PI;
        *ld = make_double2(lr, li);
    }
    __syncthreads();

    // 4. row solves x m = w_p
    for (int p = tid; p < n; p += THREADS) {
        // keep the LU factors in LDS: without this barrier the compiler hoists
        // all o^2 loads out of the row loop into registers and spills
        asm volatile("" ::: "memory");
        double2 x[MAXO];
#pragma unroll
        for (int j = 0; j < MAXO; ++j) x[j] = (j < o) ? W[p * o + j] : make_double2(0.0, 0.0);
        // u U = w
#pragma unroll
        for (int j = 0; j < MAXO; ++j) {
            if (j < o) {
                x[j] = cmul(x[j], dinv[j]);
#pragma unroll
                for (int c = j + 1; c < MAXO; ++c)
                    if (c < o) csub_mul(x[c], x[j], m[j][c]);
            }
        }
        // v L = u
#pragma unroll
        for (int j = MAXO - 1; j >= 0; --j) {
            if (j < o) {
#pragma unroll
                for (int c = 0; c < j; ++c) csub_mul(x[c], x[j], m[j][c]);
            }
        }
#pragma unroll
        for (int k = 0; k < MAXO; ++k) {
            if (k < o) {
                const int i = perm[k];
                gre[(long long)i * n + p] = x[k].x;
                gim[(long long)i * n + p] = x[k].y;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// One wave per (walker, spin), four per workgroup, for o <= 32: the same
// arithmetic as greens_body in the same order (m by the same MFMA sequence,
// the same LU updates, log det summed sequentially over k, the same row
// solves), with the LU driven by one wave in its own LDS region: no workgroup
// barriers, only the wave's LDS ordering (s_waitcnt lgkmcnt(0)).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

template <int MAXO>
struct greens_wave_lds {
    double2 m[MAXO][MAXO + 1];
    double2 dinv[MAXO];
    double2 logs[MAXO];  // (log |U_kk|, arg U_kk)
    int perm[MAXO];
};

// Optional second copy of the G rows as split-bf16 planes (afq::split_bf16):
// hi / lo of the real row and of the imaginary row, same column order.
struct split_rows {
    uint16_t *hre, *lre, *him, *lim;
};

template <int MAXO, bool SPLIT = false>
__device__ __forceinline__ void greens_wave_body(int n, int o, const double2* __restrict__ W,
                                                 const double* __restrict__ Cm, double* __restrict__ gre,
                                                 double* __restrict__ gim, double2* __restrict__ ld,
                                                 greens_wave_lds<MAXO>& S, split_rows sp = split_rows{}) {
    constexpr int NT = (MAXO + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;

    // 1. m = C^T w: every 16 x 16 tile of m by this wave, the next k-block's
    //    operands loaded before the current block's MFMAs
    {
        const int nkb = (n + 15) >> 4;
        f64x4 acc[NT][NT][2];
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) acc[ti][tj][0] = acc[ti][tj][1] = f64x4{0.0, 0.0, 0.0, 0.0};
        double a[NT][4], an[NT][4];
        double2 x[NT][4], xn[NT][4];
        auto load = [&](int kb, double (&av)[NT][4], double2 (&xv)[NT][4]) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int p = kb * 16 + lk + kk;
                const bool okp = p < n;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int c = t * 16 + lr;
                    av[t][kk] = (okp && c < o) ? Cm[p * o + c] : 0.0;
                    xv[t][kk] = (okp && c < o) ? W[p * o + c] : make_double2(0.0, 0.0);
                }
            }
        };
        load(0, a, x);
        for (int kb = 0; kb < nkb; ++kb) {
            load(kb + 1, an, xn);  // past the end: all masked to zero
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        acc[ti][tj][0] = afq::mfma_f64(a[ti][kk], x[tj][kk].x, acc[ti][tj][0]);
                        acc[ti][tj][1] = afq::mfma_f64(a[ti][kk], x[tj][kk].y, acc[ti][tj][1]);
                    }
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    a[t][kk] = an[t][kk];
                    x[t][kk] = xn[t][kk];
                }
        }
#pragma unroll
        for (int ti = 0; ti < NT; ++ti)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ii = ti * 16 + afq::mfma16<double>::row(lane, r);
                    const int j = tj * 16 + lr;
                    if (ii < o && j < o) S.m[ii][j] = make_double2(acc[ti][tj][0][r], acc[ti][tj][1][r]);
                }
    }
    if (lane < o) S.perm[lane] = lane;
    wave_lds_sync();

    // 2. LU with partial pivoting (|re| + |im|, first index on ties)
    int nswap = 0;
    for (int k = 0; k < o; ++k) {
        double best = -1.0;
        int bi = k;
        for (int r = k + lane; r < o; r += 64) {
            const double v = fabs(S.m[r][k].x) + fabs(S.m[r][k].y);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        const int piv = __builtin_amdgcn_readfirstlane(bi);
        if (piv != k) {
            for (int c = lane; c < o; c += 64) {
                const double2 t = S.m[k][c];
                S.m[k][c] = S.m[piv][c];
                S.m[piv][c] = t;
            }
            if (lane == 0) {
                const int t = S.perm[k];
                S.perm[k] = S.perm[piv];
                S.perm[piv] = t;
            }
            ++nswap;
        }
        wave_lds_sync();
        const double2 inv = crecip(S.m[k][k]);
        for (int r = k + 1 + lane; r < o; r += 64) S.m[r][k] = cmul(S.m[r][k], inv);
        wave_lds_sync();
        // trailing update: lane & 31 is the column, lane >> 5 the row parity
        // (o <= 32: one row of columns per half-wave, no index division)
        {
            const int c = k + 1 + (lane & 31);
            if (c < o) {
                const double2 mkc = S.m[k][c];
                for (int r = k + 1 + (lane >> 5); r < o; r += 2) {
                    double2 y = S.m[r][c];
                    csub_mul(y, S.m[r][k], mkc);
                    S.m[r][c] = y;
                }
            }
        }
        wave_lds_sync();
    }

    // 3. log det (summed in k order) and pivot reciprocals
    if (lane < o) {
        const double2 d = S.m[lane][lane];
        S.dinv[lane] = crecip(d);
        S.logs[lane] = make_double2(log(hypot(d.x, d.y)), atan2(d.y, d.x));
    }
    wave_lds_sync();
    if (lane == 0) {
        double lre = 0.0, lim = 0.0;
        for (int k = 0; k < o; ++k) {
            lre += S.logs[k].x;
            lim += S.logs[k].y;
        }
        if (nswap & 1) lim += M_PI;
        *ld = make_double2(lre, lim);
    }

    // 4. row solves x m = w_p, one row per lane at a time
    for (int p = lane; p < n; p += 64) {
        asm volatile("" ::: "memory");
        double2 xr[MAXO];
#pragma unroll
        for (int j = 0; j < MAXO; ++j) xr[j] = (j < o) ? W[p * o + j] : make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < MAXO; ++j) {
            if (j < o) {
                xr[j] = cmul(xr[j], S.dinv[j]);
#pragma unroll
                for (int c = j + 1; c < MAXO; ++c)
                    if (c < o) csub_mul(xr[c], xr[j], S.m[j][c]);
            }
        }
#pragma unroll
        for (int j = MAXO - 1; j >= 0; --j) {
            if (j < o) {
#pragma unroll
                for (int c = 0; c < j; ++c) csub_mul(xr[c], xr[j], S.m[j][c]);
            }
        }
#pragma unroll
        for (int k = 0; k < MAXO; ++k) {
            if (k < o) {
                const int i = S.perm[k];
                const long long e = (long long)i * n + p;
                gre[e] = xr[k].x;
                gim[e] = xr[k].y;
                if constexpr (SPLIT) {
                    afq::split_bf16(xr[k].x, sp.hre[e], sp.lre[e]);
                    afq::split_bf16(xr[k].y, sp.him[e], sp.lim[e]);
                }
            }
        }
    }
}

// SPLIT: the G rows also go to split-bf16 planes hi / lo (2 nw rows of ldk,
// real rows then imaginary rows, like greens), columns ldg..ldk-1 zeroed.
template <int MAXO, bool SPLIT = false>
__global__ __launch_bounds__(THREADS) void greens_wave_kernel(
    int nw, int n, int o, int nspin, const double2* __restrict__ wup,
    const double2* __restrict__ wdn, const double* __restrict__ cup, const double* __restrict__ cdn,
    double* __restrict__ greens, double2* __restrict__ logdet, uint16_t* __restrict__ ghi = nullptr,
    uint16_t* __restrict__ glo = nullptr, long long ldk = 0) {
    __shared__ greens_wave_lds<MAXO> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = blockIdx.x * (THREADS / 64) + wave;
    if (item >= nw * nspin) return;
    const int w = item / nspin;
    const int s = item - w * nspin;
    const long long ldg = (long long)nspin * o * n;
    double* gre = greens + (long long)w * ldg + (long long)s * o * n;
    split_rows sp{};
    if constexpr (SPLIT) {
        const long long rre = (long long)w * ldk + (long long)s * o * n;
        const long long rim = rre + (long long)nw * ldk;
        sp = split_rows{ghi + rre, glo + rre, ghi + rim, glo + rim};
        const int lane = threadIdx.x & 63;
        if (s == nspin - 1 && lane < ldk - ldg) {
            const long long c = (long long)w * ldk + ldg + lane;
            ghi[c] = 0;
            glo[c] = 0;
            ghi[c + (long long)nw * ldk] = 0;
            glo[c + (long long)nw * ldk] = 0;
        }
    }
    greens_wave_body<MAXO, SPLIT>(n, o, (s == 0 ? wup : wdn) + (long long)w * n * o, s == 0 ? cup : cdn, gre,
                                  gre + (long long)nw * ldg, logdet + (long long)w * nspin + s, L[wave], sp);
}

template <int MAXO>
__global__ __launch_bounds__(THREADS) void greens_kernel(
    int nw, int n, int o, int nspin, const double2* __restrict__ wup,
    const double2* __restrict__ wdn, const double* __restrict__ cup, const double* __restrict__ cdn,
    double* __restrict__ greens, double2* __restrict__ logdet) {
    const int b = blockIdx.x;
    const int w = b / nspin;
    const int s = b - w * nspin;
    const long long ldg = (long long)nspin * o * n;
    double* gre = greens + (long long)w * ldg + (long long)s * o * n;
    greens_body<MAXO>(n, o, (s == 0 ? wup : wdn) + (long long)w * n * o, s == 0 ? cup : cdn, gre,
                      gre + (long long)nw * ldg, logdet + (long long)w * nspin + s);
}

// One wave per walker of every system (four per workgroup): Green's function
// and log det of the system's walkers_out.
template <int MAXO>
__global__ __launch_bounds__(THREADS) void greens_grouped_kernel(int count, int total,
                                                                 const afq_system_step* __restrict__ S) {
    __shared__ greens_wave_lds<MAXO> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int b = blockIdx.x * (THREADS / 64) + wave;
    if (b >= total) return;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    const int w = b - d.walker0;
    const int n = d.n, o = d.o;
    double* gre = d.greens + (long long)w * o * n;
    greens_wave_body<MAXO>(n, o, (const double2*)d.walkers_out + (long long)w * n * o, d.trial, gre,
                           gre + (long long)d.nw * o * n, (double2*)d.logdet + w, L[wave]);
}

// the wave kernel for capacity MAXO
template <int MAXO, bool SPLIT>
void launch_greens_wave(dim3 grid, hipStream_t st, int nw, int n, int o, int nspin, const double2* wu,
                        const double2* wd, const double* cu, const double* cd, double* greens, double2* ld,
                        uint16_t* hi, uint16_t* lo, long long ldk) {
    hipLaunchKernelGGL((greens_wave_kernel<MAXO, SPLIT>), grid, dim3(THREADS), 0, st, nw, n, o, nspin, wu, wd, cu, cd,
                       greens, ld, hi, lo, ldk);
}

}  // namespace

extern "C" int afq_overlap_greens(int nw, int n, int o, int nspin, const double* walkers_up,
                                  const double* walkers_dn, const double* trial_up,
                                  const double* trial_dn, double* greens, double* logdet,
                                  void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_overlap_greens: bad sizes nw=%d n=%d o=%d", nw, n, o);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "afq_overlap_greens: nspin must be 1 or 2");
    AFQ_REQUIRE(walkers_up && trial_up && greens && logdet, "afq_overlap_greens: null pointer");
    AFQ_REQUIRE(nspin == 1 || (walkers_dn && trial_dn), "afq_overlap_greens: null spin-down pointer");
    AFQ_REQUIRE(o <= 64, "afq_overlap_greens: o=%d exceeds the supported maximum of 64", o);
    const dim3 grid((unsigned)(nw * nspin));
    hipStream_t st = (hipStream_t)stream;
    const double2* wu = (const double2*)walkers_up;
    const double2* wd = (const double2*)(nspin == 2 ? walkers_dn : walkers_up);
    const double* cd = nspin == 2 ? trial_dn : trial_up;
    double2* ld = (double2*)logdet;
    const dim3 wgrid((unsigned)afq::ceil_div(nw * nspin, THREADS / 64));
    if (o <= 8)
        launch_greens_wave<8, false>(wgrid, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld, nullptr, nullptr, 0LL);
    else if (o <= 16)
        launch_greens_wave<16, false>(wgrid, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld, nullptr, nullptr, 0LL);
    else if (o <= 32)
        launch_greens_wave<32, false>(wgrid, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld, nullptr, nullptr, 0LL);
    else
        hipLaunchKernelGGL(greens_kernel<64>, grid, dim3(THREADS), 0, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld);
    return afq::check_launch("afq_overlap_greens");
}

extern "C" int afq_overlap_greens_grouped(int count, const afq_system_step* sys, int total_walkers,
                                          int max_o, void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_overlap_greens_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys, "afq_overlap_greens_grouped: null pointer");
    AFQ_REQUIRE(max_o > 0 && max_o <= 32, "afq_overlap_greens_grouped: max_o=%d not in 1..32", max_o);
    const dim3 grid((unsigned)afq::ceil_div(total_walkers, THREADS / 64));
    hipStream_t st = (hipStream_t)stream;
    if (max_o <= 8)
        hipLaunchKernelGGL(greens_grouped_kernel<8>, grid, dim3(THREADS), 0, st, count, total_walkers, sys);
    else if (max_o <= 16)
        hipLaunchKernelGGL(greens_grouped_kernel<16>, grid, dim3(THREADS), 0, st, count, total_walkers, sys);
    else
        hipLaunchKernelGGL(greens_grouped_kernel<32>, grid, dim3(THREADS), 0, st, count, total_walkers, sys);
    return afq::check_launch("afq_overlap_greens_grouped");
}

extern "C" int afq_overlap_greens_split(int nw, int n, int o, int nspin, const double* walkers_up,
                                        const double* walkers_dn, const double* trial_up,
                                        const double* trial_dn, double* greens, double* logdet,
                                        unsigned short* g_hi, unsigned short* g_lo, long long ldk,
                                        void* stream) {
    const long long ldg = (long long)nspin * o * n;
    AFQ_REQUIRE(g_hi && g_lo, "afq_overlap_greens_split: null pointer");
    AFQ_REQUIRE(ldk >= ldg && ldk - ldg < 64, "afq_overlap_greens_split: ldk=%lld not in [%lld, %lld)", ldk, ldg,
                ldg + 64);
    if (o > 32) {
        // block kernel, then the split of its rows
        const int rc = afq_overlap_greens(nw, n, o, nspin, walkers_up, walkers_dn, trial_up, trial_dn, greens,
                                          logdet, stream);
        if (rc != 0) return rc;
        return afq_split_bf16(2 * nw, (int)ldg, greens, ldg, g_hi, g_lo, ldk, stream);
    }
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_overlap_greens_split: bad sizes nw=%d n=%d o=%d", nw, n,
                o);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "afq_overlap_greens_split: nspin must be 1 or 2");
    AFQ_REQUIRE(walkers_up && trial_up && greens && logdet, "afq_overlap_greens_split: null pointer");
    AFQ_REQUIRE(nspin == 1 || (walkers_dn && trial_dn), "afq_overlap_greens_split: null spin-down pointer");
    hipStream_t st = (hipStream_t)stream;
    const double2* wu = (const double2*)walkers_up;
    const double2* wd = (const double2*)(nspin == 2 ? walkers_dn : walkers_up);
    const double* cd = nspin == 2 ? trial_dn : trial_up;
    double2* ld = (double2*)logdet;
    uint16_t* hi = (uint16_t*)g_hi;
    uint16_t* lo = (uint16_t*)g_lo;
    const dim3 wgrid((unsigned)afq::ceil_div(nw * nspin, THREADS / 64));
    if (o <= 8)
        launch_greens_wave<8, true>(wgrid, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld, hi, lo, ldk);
    else if (o <= 16)
        launch_greens_wave<16, true>(wgrid, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld, hi, lo, ldk);
    else
        launch_greens_wave<32, true>(wgrid, st, nw, n, o, nspin, wu, wd, trial_up, cd, greens, ld, hi, lo, ldk);
    return afq::check_launch("afq_overlap_greens_split");
}
