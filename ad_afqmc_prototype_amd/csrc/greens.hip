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
// CT: complex trial orbitals, Cm points at interleaved (re, im) pairs and
// m = C^H w = sum_p (Cr - i Ci)(wr + i wi).
template <int MAXO, bool CT = false>
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
                    const double2 x = (okp && j < o) ? W[p * o + j] : make_double2(0.0, 0.0);
                    if constexpr (CT) {
                        const double2 a = (okp && i < o) ? reinterpret_cast<const double2*>(Cm)[p * o + i]
                                                         : make_double2(0.0, 0.0);
                        ar = afq::mfma_f64(a.x, x.x, ar);
                        ar = afq::mfma_f64(a.y, x.y, ar);
                        ai = afq::mfma_f64(a.x, x.y, ai);
                        ai = afq::mfma_f64(-a.y, x.x, ai);
                    } else {
                        const double a = (okp && i < o) ? Cm[p * o + i] : 0.0;
                        ar = afq::mfma_f64(a, x.x, ar);
                        ai = afq::mfma_f64(a, x.y, ai);
                    }
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

template <int MAXO, bool SPLIT = false, bool CT = false>
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
        // complex trial (CT): Cm holds interleaved (re, im) pairs and
        // m = C^H w = sum_p (Cr - i Ci)(wr + i wi)
        constexpr int NA = CT ? 2 : 1;
        double a[NT][4][NA], an[NT][4][NA];
        double2 x[NT][4], xn[NT][4];
        auto load = [&](int kb, double (&av)[NT][4][NA], double2 (&xv)[NT][4]) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int p = kb * 16 + lk + kk;
                const bool okp = p < n;
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    const int c = t * 16 + lr;
                    if constexpr (CT) {
                        const double2 cv = (okp && c < o) ? reinterpret_cast<const double2*>(Cm)[p * o + c]
                                                          : make_double2(0.0, 0.0);
                        av[t][kk][0] = cv.x;
                        av[t][kk][NA - 1] = cv.y;
                    } else {
                        av[t][kk][0] = (okp && c < o) ? Cm[p * o + c] : 0.0;
                    }
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
                        acc[ti][tj][0] = afq::mfma_f64(a[ti][kk][0], x[tj][kk].x, acc[ti][tj][0]);
                        acc[ti][tj][1] = afq::mfma_f64(a[ti][kk][0], x[tj][kk].y, acc[ti][tj][1]);
                        if constexpr (CT) {
                            acc[ti][tj][0] = afq::mfma_f64(a[ti][kk][NA - 1], x[tj][kk].y, acc[ti][tj][0]);
                            acc[ti][tj][1] = afq::mfma_f64(-a[ti][kk][NA - 1], x[tj][kk].x, acc[ti][tj][1]);
                        }
                    }
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
#pragma unroll
                    for (int e = 0; e < NA; ++e) a[t][kk][e] = an[t][kk][e];
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

// ---------------------------------------------------------------------------
// Split-bf16 Green's function of the bf16x3 step, o <= 16 NT: the rows of G
// only ever feed the split-bf16 force-bias GEMM there, so they are formed as a
// product on the bf16 matrix cores and written straight as its bf16 planes.
//   1. m = C^T w on v_mfma_f64_16x16x4 (fp64, as greens_wave_body);
//   2. Gauss-Jordan inversion with partial pivoting (|re| + |im|, first index
//      on ties: the LU's pivots) in the m accumulator registers, lane
//      (lr, q) holding a[s][t] = m[4 s + q][16 t + lr]; the pivots give
//      log det m in fp64 exactly as the LU does;
//   3. A = (P m)^-1 split hi / lo into LDS as A^T, then G[perm[j]][p] =
//      sum_k A[k][j] w[p][k] on v_mfma_f32_16x16x32_bf16 (al.bh + ah.bl +
//      ah.bh per real product), each 16-row tile of w split on load.
// One wave per (walker, spin), four per workgroup; broadcasts between lanes
// go through the wave's own LDS region (no workgroup barriers).
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

template <int NT>
struct gj_lds {
    static constexpr int O = 16 * NT;
    static constexpr int AP = O + 8;  // bf16 pitch of one row of A^T (16-byte multiple)
    double2 col[O];
    double2 rowk[O];
    double2 rowp[O];
    double2 piv[O];  // the pivots, for the log det after the elimination
    int perm[O];
    uint16_t at[4][O][AP];  // A^T planes: re hi, re lo, im hi, im lo
};

__device__ __forceinline__ bf16x8_t neg_bf16x8(bf16x8_t x) {
    uint4 u = __builtin_bit_cast(uint4, x);
    u.x ^= 0x80008000u;
    u.y ^= 0x80008000u;
    u.z ^= 0x80008000u;
    u.w ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8_t, u);
}

// lane src's value of x (src uniform)
__device__ __forceinline__ double readlane_f64(double x, int src) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, src);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), src);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__device__ __forceinline__ f32x4 mfma_bf16x32(bf16x8_t a, bf16x8_t b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int NT>
__device__ __forceinline__ void greens_gj_split_body(int n, int o, const double2* __restrict__ W,
                                                     const double* __restrict__ Cm, double2* __restrict__ ld,
                                                     gj_lds<NT>& S, split_rows sp) {
    constexpr int O = 16 * NT;
    constexpr int NS = 4 * NT;  // row slots per lane
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15;
    const int q = lane >> 4;
    const int lk = q * 4;

    // ---- 1. m = C^T w (fp64 MFMA, next k-block's operands loaded ahead) ----
    f64x4 acc[NT][NT][2];
    {
        const int nkb = (n + 15) >> 4;
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
            load(kb + 1, an, xn);
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
    }
    // accumulator element r of tile (ti, tj): m[16 ti + 4 r + q][16 tj + lr]
    double2 m[NS][NT];
#pragma unroll
    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) m[4 * ti + r][tj] = make_double2(acc[ti][tj][0][r], acc[ti][tj][1][r]);
    if (lane < O) S.perm[lane] = lane;

    // ---- 2. Gauss-Jordan with partial pivoting ----
    int nswap = 0;
    // k = 4 ks + kq: the row slot ks (and the column tile kt) are compile-time
    // constants of the unrolled outer loop, so every register index is static
#pragma unroll
    for (int ks = 0; ks < NS; ++ks) {
        const int kt = ks >> 2;
#pragma unroll 1
        for (int kq = 0; kq < 4; ++kq) {
            const int k = 4 * ks + kq;
            if (k >= o) break;
            const int kl = k & 15;
            // pivot: rows k..o-1 of column k (lanes lr == kl hold it)
            double best = -1.0;
            int bi = k;
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) {
                const int r = 4 * s2 + q;
                const double v = fabs(m[s2][kt].x) + fabs(m[s2][kt].y);
                if (r >= k && r < o && v > best) {
                    best = v;
                    bi = r;
                }
            }
#pragma unroll
            for (int off = 16; off < 64; off <<= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            const int piv = __builtin_amdgcn_readlane(bi, kl);
            if (piv != k) {
                const int pq = piv & 3, ps = piv >> 2;
                // rows k and piv through LDS
                if (q == kq) {
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) S.rowk[tj * 16 + lr] = m[ks][tj];
                }
                if (q == pq) {
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        double2 v = m[0][tj];
#pragma unroll
                        for (int s2 = 1; s2 < NS; ++s2)
                            if (s2 == ps) v = m[s2][tj];
                        S.rowp[tj * 16 + lr] = v;
                    }
                }
                if (lane == 0) {
                    const int t = S.perm[k];
                    S.perm[k] = S.perm[piv];
                    S.perm[piv] = t;
                }
                wave_lds_sync();
                if (q == kq) {
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) m[ks][tj] = S.rowp[tj * 16 + lr];
                }
                if (q == pq) {
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        const double2 v = S.rowk[tj * 16 + lr];
#pragma unroll
                        for (int s2 = 0; s2 < NS; ++s2)
                            if (s2 == ps) m[s2][tj] = v;
                    }
                }
                ++nswap;
                wave_lds_sync();
            }
            // pivot value (uniform): lane (kl, kq), slot (ks, kt)
            const int src = kl + 16 * kq;
            const double2 pv = make_double2(readlane_f64(m[ks][kt].x, src), readlane_f64(m[ks][kt].y, src));
            if (lane == 0) S.piv[k] = pv;
            const double2 inv = crecip(pv);
            // scale row k (A[k][k] <- 1 / p) and publish it and column k
            if (q == kq) {
#pragma unroll
                for (int tj = 0; tj < NT; ++tj) {
                    const double2 v = (tj * 16 + lr == k) ? inv : cmul(m[ks][tj], inv);
                    m[ks][tj] = v;
                    S.rowk[tj * 16 + lr] = v;
                }
            }
            if (lr == kl) {
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) S.col[4 * s2 + q] = m[s2][kt];
            }
            wave_lds_sync();
            double2 rk[NT];
#pragma unroll
            for (int tj = 0; tj < NT; ++tj) rk[tj] = S.rowk[tj * 16 + lr];
#pragma unroll
            for (int s2 = 0; s2 < NS; ++s2) {
                const int r = 4 * s2 + q;
                if (r != k) {
                    const double2 f = S.col[r];
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        double2 y = (tj * 16 + lr == k) ? make_double2(0.0, 0.0) : m[s2][tj];
                        csub_mul(y, f, rk[tj]);
                        m[s2][tj] = y;
                    }
                }
            }
            wave_lds_sync();
        }
    }
    // log det = sum_k log |p_k| + i arg p_k (+ i pi per swap): lane k evaluates
    // pivot k's logarithms (off the elimination's latency chain), the sums run
    // over k in order
    double lre = 0.0, lim = 0.0;
    {
        wave_lds_sync();
        double ll = 0.0, la = 0.0;
        if (lane < o) {
            const double2 pv = S.piv[lane];
            ll = log(hypot(pv.x, pv.y));
            la = atan2(pv.y, pv.x);
        }
        for (int k = 0; k < o; ++k) {
            lre += readlane_f64(ll, k);
            lim += readlane_f64(la, k);
        }
    }
    if (nswap & 1) lim += M_PI;
    if (lane == 0) *ld = make_double2(lre, lim);

    // ---- 3. A^T (split) into LDS, then G rows = A^T w^T on bf16 MFMA ----
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2)
#pragma unroll
        for (int tj = 0; tj < NT; ++tj) {
            const int j = tj * 16 + lr, kk = 4 * s2 + q;
            uint16_t h, l;
            afq::split_bf16(m[s2][tj].x, h, l);
            S.at[0][j][kk] = h;
            S.at[1][j][kk] = l;
            afq::split_bf16(m[s2][tj].y, h, l);
            S.at[2][j][kk] = h;
            S.at[3][j][kk] = l;
        }
    wave_lds_sync();
    // A fragments of output tile tjo, k-block kb (32 wide): row j = 16 tjo + lr,
    // k = 32 kb + 8 q .. + 7
    constexpr int NKB = (O + 31) / 32;
    bf16x8_t Arh[NT][NKB], Arl[NT][NKB], Aih[NT][NKB], Ail[NT][NKB];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            const int kk = kb * 32 + 8 * q;
            if (kk < O) {
                Arh[t][kb] = *reinterpret_cast<const bf16x8_t*>(&S.at[0][t * 16 + lr][kk]);
                Arl[t][kb] = *reinterpret_cast<const bf16x8_t*>(&S.at[1][t * 16 + lr][kk]);
                Aih[t][kb] = *reinterpret_cast<const bf16x8_t*>(&S.at[2][t * 16 + lr][kk]);
                Ail[t][kb] = *reinterpret_cast<const bf16x8_t*>(&S.at[3][t * 16 + lr][kk]);
            } else {
                Arh[t][kb] = Arl[t][kb] = Aih[t][kb] = Ail[t][kb] = bf16x8_t{};
            }
        }
    int prow[NT][4];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = t * 16 + 4 * q + r;
            prow[t][r] = j < o ? S.perm[j] : -1;
        }
    // w rows p = 16 pt + lr, columns 8 q .. 8 q + 7 of each k-block, as fp32
    auto load_w = [&](int pt, float (&re)[NKB][8], float (&im)[NKB][8]) {
        const int p = pt * 16 + lr;
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int kk = kb * 32 + 8 * q + e;
                const double2 v = (p < n && kk < o) ? W[p * o + kk] : make_double2(0.0, 0.0);
                re[kb][e] = (float)v.x;
                im[kb][e] = (float)v.y;
            }
    };
    auto split8 = [&](const float (&x)[8], bf16x8_t& h, bf16x8_t& l) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const __bf16 hh = (__bf16)x[e];
            h[e] = hh;
            l[e] = (__bf16)(x[e] - (float)hh);
        }
    };
    // G rows of w rows 16 pt .. 16 pt + 15 (w as fp32, the lane's fragment
    // entries): split, the product on bf16 MFMA, the bf16 planes stored
    auto row_block = [&](int pt, const float (&wre)[NKB][8], const float (&wim)[NKB][8]) {
        bf16x8_t brh[NKB], brl[NKB], bih[NKB], bil[NKB];
#pragma unroll
        for (int kb = 0; kb < NKB; ++kb) {
            split8(wre[kb], brh[kb], brl[kb]);
            split8(wim[kb], bih[kb], bil[kb]);
        }
        const int p = pt * 16 + lr;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < NKB; ++kb) {
                // re = Ar Wr - Ai Wi, im = Ar Wi + Ai Wr
                cr = mfma_bf16x32(Arl[t][kb], brh[kb], cr);
                cr = mfma_bf16x32(Arh[t][kb], brl[kb], cr);
                cr = mfma_bf16x32(Arh[t][kb], brh[kb], cr);
                ci = mfma_bf16x32(Arl[t][kb], bih[kb], ci);
                ci = mfma_bf16x32(Arh[t][kb], bil[kb], ci);
                ci = mfma_bf16x32(Arh[t][kb], bih[kb], ci);
                const bf16x8_t nih = neg_bf16x8(Aih[t][kb]), nil = neg_bf16x8(Ail[t][kb]);
                cr = mfma_bf16x32(nil, bih[kb], cr);
                cr = mfma_bf16x32(nih, bil[kb], cr);
                cr = mfma_bf16x32(nih, bih[kb], cr);
                ci = mfma_bf16x32(Ail[t][kb], brh[kb], ci);
                ci = mfma_bf16x32(Aih[t][kb], brl[kb], ci);
                ci = mfma_bf16x32(Aih[t][kb], brh[kb], ci);
            }
            if (p < n) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int i = prow[t][r];
                    if (i >= 0) {
                        const long long e = (long long)i * n + p;
                        const __bf16 hr = (__bf16)cr[r];
                        const __bf16 hi_ = (__bf16)ci[r];
                        sp.hre[e] = __builtin_bit_cast(uint16_t, hr);
                        sp.lre[e] = __builtin_bit_cast(uint16_t, (__bf16)(cr[r] - (float)hr));
                        sp.him[e] = __builtin_bit_cast(uint16_t, hi_);
                        sp.lim[e] = __builtin_bit_cast(uint16_t, (__bf16)(ci[r] - (float)hi_));
                    }
                }
            }
        }
    };
    const int npt = (n + 15) >> 4;
    if constexpr (NT == 2) {
        // 16 rows of w per step: one contiguous block of 16 o entries, read
        // coalesced into registers one step ahead and passed to the fragment
        // layout through the A^T region (free once A's fragments are loaded):
        // each load instruction then covers 8 cache lines, not 64
        constexpr int PW = O + 1;        // padded row pitch (double2)
        constexpr int NV = 16 * O / 64;  // entries per lane per block
        static_assert(16 * PW * sizeof(double2) <= sizeof(S.at), "w block exceeds the A^T region");
        double2* Wt = reinterpret_cast<double2*>(&S.at[0][0][0]);
        wave_lds_sync();  // A's fragments and the permutation are read before the region is reused
        int soff[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int u = i * 64 + lane;
            const int row = u / o;
            soff[i] = u < 16 * o ? row * PW + (u - row * o) : -1;
        }
        double2 wv[NV];
        auto fetch = [&](int pt) {
            const int base = pt * 16 * o;
            const int lim = min(16, n - pt * 16) * o;  // <= 0 past the end: nothing read
#pragma unroll
            for (int i = 0; i < NV; ++i) {
                const int u = i * 64 + lane;
                wv[i] = u < lim ? W[base + u] : make_double2(0.0, 0.0);
            }
        };
        fetch(0);
        for (int pt = 0; pt < npt; ++pt) {
#pragma unroll
            for (int i = 0; i < NV; ++i)
                if (soff[i] >= 0) Wt[soff[i]] = wv[i];
            fetch(pt + 1);
            wave_lds_sync();
            float wre[NKB][8], wim[NKB][8];
#pragma unroll
            for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int kk = kb * 32 + 8 * q + e;
                    const double2 v = kk < o ? Wt[lr * PW + kk] : make_double2(0.0, 0.0);
                    wre[kb][e] = (float)v.x;
                    wim[kb][e] = (float)v.y;
                }
            wave_lds_sync();  // the block is read before the next one is written
            row_block(pt, wre, wim);
        }
    } else {
        float wre[NKB][8], wim[NKB][8], nre[NKB][8], nim[NKB][8];
        load_w(0, wre, wim);
        for (int pt = 0; pt < npt; ++pt) {
            load_w(pt + 1, nre, nim);  // past the end: all zeros, nothing read
            row_block(pt, wre, wim);
#pragma unroll
            for (int kb = 0; kb < NKB; ++kb)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    wre[kb][e] = nre[kb][e];
                    wim[kb][e] = nim[kb][e];
                }
        }
    }
}

template <int NT>
__global__ __launch_bounds__(THREADS) void greens_gj_split_kernel(
    int nw, int n, int o, int nspin, const double2* __restrict__ wup, const double2* __restrict__ wdn,
    const double* __restrict__ cup, const double* __restrict__ cdn, double2* __restrict__ logdet,
    uint16_t* __restrict__ ghi, uint16_t* __restrict__ glo, long long ldk) {
    __shared__ gj_lds<NT> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = blockIdx.x * (THREADS / 64) + wave;
    if (item >= nw * nspin) return;
    const int w = item / nspin;
    const int s = item - w * nspin;
    const long long ldg = (long long)nspin * o * n;
    const long long rre = (long long)w * ldk + (long long)s * o * n;
    const long long rim = rre + (long long)nw * ldk;
    const int lane = threadIdx.x & 63;
    if (s == nspin - 1 && lane < ldk - ldg) {
        const long long c = (long long)w * ldk + ldg + lane;
        ghi[c] = 0;
        glo[c] = 0;
        ghi[c + (long long)nw * ldk] = 0;
        glo[c + (long long)nw * ldk] = 0;
    }
    greens_gj_split_body<NT>(n, o, (s == 0 ? wup : wdn) + (long long)w * n * o, s == 0 ? cup : cdn,
                             logdet + (long long)w * nspin + s, L[wave],
                             split_rows{ghi + rre, glo + rre, ghi + rim, glo + rim});
}

// Every system's walkers (restricted), one wave each: wave item -> (system,
// walker) by the walker0 prefix sums; planes and pitch from the descriptor.
template <int NT>
__global__ __launch_bounds__(THREADS) void greens_gj_split_grouped_kernel(int count, int total,
                                                                         const afq_system_step* __restrict__ S) {
    __shared__ gj_lds<NT> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = blockIdx.x * (THREADS / 64) + wave;
    if (item >= total) return;
    const int si = afq::find_group(count, item, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    const int w = item - d.walker0;
    const int n = d.n, o = d.o;
    const long long ldk = d.ldk;
    const long long ldg = (long long)o * n;
    uint16_t* ghi = d.greens_hi;
    uint16_t* glo = d.greens_lo;
    const long long rre = (long long)w * ldk;
    const long long rim = rre + (long long)d.nw * ldk;
    const int lane = threadIdx.x & 63;
    if (lane < ldk - ldg) {
        ghi[rre + ldg + lane] = 0;
        glo[rre + ldg + lane] = 0;
        ghi[rim + ldg + lane] = 0;
        glo[rim + ldg + lane] = 0;
    }
    greens_gj_split_body<NT>(n, o, (const double2*)d.walkers_out + (long long)w * n * o, d.trial,
                             (double2*)d.logdet + w, L[wave], split_rows{ghi + rre, glo + rre, ghi + rim, glo + rim});
}

// Complex trial orbitals (interleaved complex128 C), fp64 rows only.
template <int MAXO>
__global__ __launch_bounds__(THREADS) void greens_wave_ctrial_kernel(
    int nw, int n, int o, int nspin, const double2* __restrict__ wup, const double2* __restrict__ wdn,
    const double* __restrict__ cup, const double* __restrict__ cdn, double* __restrict__ greens,
    double2* __restrict__ logdet) {
    __shared__ greens_wave_lds<MAXO> L[THREADS / 64];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = blockIdx.x * (THREADS / 64) + wave;
    if (item >= nw * nspin) return;
    const int w = item / nspin;
    const int s = item - w * nspin;
    const long long ldg = (long long)nspin * o * n;
    double* gre = greens + (long long)w * ldg + (long long)s * o * n;
    greens_wave_body<MAXO, false, true>(n, o, (s == 0 ? wup : wdn) + (long long)w * n * o, s == 0 ? cup : cdn, gre,
                                        gre + (long long)nw * ldg, logdet + (long long)w * nspin + s, L[wave]);
}

template <int MAXO>
__global__ __launch_bounds__(THREADS) void greens_ctrial_kernel(
    int nw, int n, int o, int nspin, const double2* __restrict__ wup,
    const double2* __restrict__ wdn, const double* __restrict__ cup, const double* __restrict__ cdn,
    double* __restrict__ greens, double2* __restrict__ logdet) {
    const int b = blockIdx.x;
    const int w = b / nspin;
    const int s = b - w * nspin;
    const long long ldg = (long long)nspin * o * n;
    double* gre = greens + (long long)w * ldg + (long long)s * o * n;
    greens_body<MAXO, true>(n, o, (s == 0 ? wup : wdn) + (long long)w * n * o, s == 0 ? cup : cdn, gre,
                            gre + (long long)nw * ldg, logdet + (long long)w * nspin + s);
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

// ---------------------------------------------------------------------------
// Adjoint of the Green's function and log-overlap (meas/rhf.py:15-24,
// trial/rhf.py:25-44 differentiated; restricted walkers, real C): for
// G = (w m^-1)^T, ld = log det m, m = C^T w, and complex adjoints G_bar
// (o x n, optional) and ld_bar (optional),
//   w_bar = [G_bar^T - C (m^-H Z - ld_bar I)] m^-H,   Z = w^H G_bar^T.
// One 256-thread workgroup per walker: m and Z by direct sums over the n
// orbitals, m^-1 by Gauss-Jordan elimination with partial pivoting on the
// augmented [m | I] in LDS, then the two o-deep products.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(THREADS) void greens_adjoint_kernel(int n, int o, const double2* __restrict__ W,
                                                                 const double* __restrict__ Cm,
                                                                 const double2* __restrict__ gbar,
                                                                 const double2* __restrict__ ldbar,
                                                                 double2* __restrict__ wbar, int accumulate) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int o2 = 2 * o;
    double2* aug = reinterpret_cast<double2*>(lds_raw);  // [o][2o]: [m | I] -> [I | m^-1]
    double2* Z = aug + o * o2;                            // [o][o]: Z
    double2* Q = Z + o * o;                               // [o][o]: m^-H Z - ld_bar I
    double2* Y = Q + o * o;                               // [n][o]
    __shared__ double2 colk[64];
    __shared__ int piv;
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const double2* Ww = W + (long long)w * n * o;
    const double2* Gb = gbar ? gbar + (long long)w * o * n : nullptr;
    const double2 lb = ldbar ? ldbar[w] : make_double2(0.0, 0.0);

    // m = C^T w, Z = w^H G_bar^T
    for (int e = tid; e < o * o; e += THREADS) {
        const int i = e / o, j = e - (e / o) * o;
        double2 m = make_double2(0.0, 0.0), z = make_double2(0.0, 0.0);
        for (int p = 0; p < n; ++p) {
            const double2 wj = Ww[p * o + j];
            const double c = Cm[p * o + i];
            m.x = fma(c, wj.x, m.x);
            m.y = fma(c, wj.y, m.y);
            if (Gb) {
                const double2 wi = Ww[p * o + i];
                const double2 g = Gb[j * n + p];
                // conj(wi) g
                z.x = fma(wi.x, g.x, fma(wi.y, g.y, z.x));
                z.y = fma(wi.x, g.y, fma(-wi.y, g.x, z.y));
            }
        }
        aug[i * o2 + j] = m;
        aug[i * o2 + o + j] = make_double2(i == j ? 1.0 : 0.0, 0.0);
        Z[e] = z;
    }
    __syncthreads();
    // Gauss-Jordan with partial pivoting
    for (int k = 0; k < o; ++k) {
        if (tid < 64) {
            // pivot: the largest |aug[r][k]|, r >= k (lowest r on ties)
            double best = -1.0;
            int br = k;
            for (int r = k + tid; r < o; r += 64) {
                const double2 v = aug[r * o2 + k];
                const double a = v.x * v.x + v.y * v.y;
                if (a > best) {
                    best = a;
                    br = r;
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const int orr = __shfl_xor(br, off, 64);
                if (ob > best || (ob == best && orr < br)) {
                    best = ob;
                    br = orr;
                }
            }
            if (tid == 0) piv = br;
        }
        __syncthreads();
        const int r = piv;
        if (r != k) {
            for (int j = tid; j < o2; j += THREADS) {
                const double2 t = aug[k * o2 + j];
                aug[k * o2 + j] = aug[r * o2 + j];
                aug[r * o2 + j] = t;
            }
        }
        __syncthreads();
        const double2 inv = crecip(aug[k * o2 + k]);
        if (tid < o) colk[tid] = aug[tid * o2 + k];  // the multipliers, before the row is scaled
        __syncthreads();
        for (int j = tid; j < o2; j += THREADS) aug[k * o2 + j] = cmul(aug[k * o2 + j], inv);
        __syncthreads();
        for (int e = tid; e < o * o2; e += THREADS) {
            const int i = e / o2, j = e - (e / o2) * o2;
            if (i != k) csub_mul(aug[e], colk[i], aug[k * o2 + j]);
        }
        __syncthreads();
    }
    // A = m^-1 = aug[:, o:]; Q = A^H Z - ld_bar I
    for (int e = tid; e < o * o; e += THREADS) {
        const int i = e / o, j = e - (e / o) * o;
        double2 v = make_double2(i == j ? -lb.x : 0.0, i == j ? -lb.y : 0.0);
        if (Gb) {
            for (int k = 0; k < o; ++k) {
                const double2 a = aug[k * o2 + o + i];
                const double2 z = Z[k * o + j];
                v.x = fma(a.x, z.x, fma(a.y, z.y, v.x));
                v.y = fma(a.x, z.y, fma(-a.y, z.x, v.y));
            }
        }
        Q[e] = v;
    }
    __syncthreads();
    // Y = G_bar^T - C Z
    for (int e = tid; e < n * o; e += THREADS) {
        const int p = e / o, j = e - (e / o) * o;
        double2 y = Gb ? Gb[j * n + p] : make_double2(0.0, 0.0);
        for (int i = 0; i < o; ++i) {
            const double c = Cm[p * o + i];
            const double2 z = Q[i * o + j];
            y.x = fma(-c, z.x, y.x);
            y.y = fma(-c, z.y, y.y);
        }
        Y[e] = y;
    }
    __syncthreads();
    // w_bar = Y A^H:  w_bar[p][j] = sum_i Y[p][i] conj(A[j][i])
    double2* out = wbar + (long long)w * n * o;
    for (int e = tid; e < n * o; e += THREADS) {
        const int p = e / o, j = e - (e / o) * o;
        double2 v = accumulate ? out[e] : make_double2(0.0, 0.0);
        for (int i = 0; i < o; ++i) {
            const double2 y = Y[p * o + i];
            const double2 a = aug[j * o2 + o + i];
            v.x = fma(y.x, a.x, fma(y.y, a.y, v.x));
            v.y = fma(y.y, a.x, fma(-y.x, a.y, v.y));
        }
        out[e] = v;
    }
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

extern "C" int afq_overlap_greens_bf16x3_grouped(int count, const afq_system_step* sys, int total_walkers,
                                                 int max_o, void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_overlap_greens_bf16x3_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys, "afq_overlap_greens_bf16x3_grouped: null pointer");
    AFQ_REQUIRE(max_o > 0 && max_o <= 32, "afq_overlap_greens_bf16x3_grouped: max_o=%d not in 1..32", max_o);
    const dim3 grid((unsigned)afq::ceil_div(total_walkers, THREADS / 64));
    hipStream_t st = (hipStream_t)stream;
    if (max_o <= 16)
        hipLaunchKernelGGL(greens_gj_split_grouped_kernel<1>, grid, dim3(THREADS), 0, st, count, total_walkers, sys);
    else
        hipLaunchKernelGGL(greens_gj_split_grouped_kernel<2>, grid, dim3(THREADS), 0, st, count, total_walkers, sys);
    return afq::check_launch("afq_overlap_greens_bf16x3_grouped");
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

extern "C" int afq_overlap_greens_bf16x3(int nw, int n, int o, int nspin, const double* walkers_up,
                                         const double* walkers_dn, const double* trial_up, const double* trial_dn,
                                         double* logdet, unsigned short* g_hi, unsigned short* g_lo, long long ldk,
                                         void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_overlap_greens_bf16x3: bad sizes nw=%d n=%d o=%d", nw, n,
                o);
    AFQ_REQUIRE(o <= 32, "afq_overlap_greens_bf16x3: o=%d exceeds the maximum of 32", o);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "afq_overlap_greens_bf16x3: nspin must be 1 or 2");
    AFQ_REQUIRE(walkers_up && trial_up && logdet && g_hi && g_lo, "afq_overlap_greens_bf16x3: null pointer");
    AFQ_REQUIRE(nspin == 1 || (walkers_dn && trial_dn), "afq_overlap_greens_bf16x3: null spin-down pointer");
    const long long ldg = (long long)nspin * o * n;
    AFQ_REQUIRE(ldk >= ldg && ldk - ldg < 64, "afq_overlap_greens_bf16x3: ldk=%lld not in [%lld, %lld)", ldk, ldg,
                ldg + 64);
    AFQ_REQUIRE((long long)n * o < (1LL << 31), "afq_overlap_greens_bf16x3: n * o too large");
    hipStream_t st = (hipStream_t)stream;
    const double2* wu = (const double2*)walkers_up;
    const double2* wd = (const double2*)(nspin == 2 ? walkers_dn : walkers_up);
    const double* cd = nspin == 2 ? trial_dn : trial_up;
    const dim3 grid((unsigned)afq::ceil_div(nw * nspin, THREADS / 64));
    if (o <= 16)
        hipLaunchKernelGGL(greens_gj_split_kernel<1>, grid, dim3(THREADS), 0, st, nw, n, o, nspin, wu, wd, trial_up,
                           cd, (double2*)logdet, (uint16_t*)g_hi, (uint16_t*)g_lo, ldk);
    else
        hipLaunchKernelGGL(greens_gj_split_kernel<2>, grid, dim3(THREADS), 0, st, nw, n, o, nspin, wu, wd, trial_up,
                           cd, (double2*)logdet, (uint16_t*)g_hi, (uint16_t*)g_lo, ldk);
    return afq::check_launch("afq_overlap_greens_bf16x3");
}

extern "C" int afq_overlap_greens_ctrial(int nw, int n, int o, int nspin, const double* walkers_up,
                                         const double* walkers_dn, const double* trial_up,
                                         const double* trial_dn, double* greens, double* logdet, void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_overlap_greens_ctrial: bad sizes nw=%d n=%d o=%d", nw, n,
                o);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "afq_overlap_greens_ctrial: nspin must be 1 or 2");
    AFQ_REQUIRE(walkers_up && trial_up && greens && logdet, "afq_overlap_greens_ctrial: null pointer");
    AFQ_REQUIRE(nspin == 1 || (walkers_dn && trial_dn), "afq_overlap_greens_ctrial: null spin-down pointer");
    AFQ_REQUIRE(o <= 64, "afq_overlap_greens_ctrial: o=%d exceeds the supported maximum of 64", o);
    hipStream_t st = (hipStream_t)stream;
    const double2* wu = (const double2*)walkers_up;
    const double2* wd = (const double2*)(nspin == 2 ? walkers_dn : walkers_up);
    const double* cd = nspin == 2 ? trial_dn : trial_up;
    double2* ld = (double2*)logdet;
    const dim3 wgrid((unsigned)afq::ceil_div(nw * nspin, THREADS / 64));
    if (o <= 16)
        hipLaunchKernelGGL(greens_wave_ctrial_kernel<16>, wgrid, dim3(THREADS), 0, st, nw, n, o, nspin, wu, wd,
                           trial_up, cd, greens, ld);
    else if (o <= 32)
        hipLaunchKernelGGL(greens_wave_ctrial_kernel<32>, wgrid, dim3(THREADS), 0, st, nw, n, o, nspin, wu, wd,
                           trial_up, cd, greens, ld);
    else
        hipLaunchKernelGGL(greens_ctrial_kernel<64>, dim3((unsigned)(nw * nspin)), dim3(THREADS), 0, st, nw, n, o,
                           nspin, wu, wd, trial_up, cd, greens, ld);
    return afq::check_launch("afq_overlap_greens_ctrial");
}

extern "C" int afq_greens_adjoint(int nw, int n, int o, const double* walkers, const double* trial,
                                  const double* g_bar, const double* ld_bar, double* w_bar, int accumulate,
                                  void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_greens_adjoint: bad sizes nw=%d n=%d o=%d", nw, n, o);
    AFQ_REQUIRE(o <= 32, "afq_greens_adjoint: o=%d exceeds the supported maximum of 32", o);
    AFQ_REQUIRE(walkers && trial && w_bar, "afq_greens_adjoint: null pointer");
    AFQ_REQUIRE(g_bar || ld_bar, "afq_greens_adjoint: neither g_bar nor ld_bar given");
    const size_t lds = sizeof(double2) * ((size_t)o * 2 * o + (size_t)2 * o * o + (size_t)n * o);
    AFQ_REQUIRE(lds <= 150 * 1024, "afq_greens_adjoint: n=%d o=%d needs %zu bytes of LDS", n, o, lds);
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)greens_adjoint_kernel);
    hipLaunchKernelGGL(greens_adjoint_kernel, dim3((unsigned)nw), dim3(THREADS), lds, (hipStream_t)stream, n, o,
                       (const double2*)walkers, trial, (const double2*)g_bar, (const double2*)ld_bar,
                       (double2*)w_bar, accumulate);
    return afq::check_launch("afq_greens_adjoint");
}
