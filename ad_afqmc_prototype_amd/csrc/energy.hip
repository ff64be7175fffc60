// Local energy (mixed estimator) per walker; fp64 or fp32 contractions, fp64
// accumulation of every sum over g.
//
// One 256-thread workgroup per WPB walkers.  G^T (n x o, split planes) of one
// spin sector of each of its walkers sits in LDS as the MFMA B operand; wave wv
// handles Cholesky vectors g = wv, wv+4, ...: f_g = rot_chol_g G^T (o x o) is
// accumulated in MFMA registers from rot_chol_g, streamed from L2/MALL once per
// workgroup and used for all WPB walkers, then reduced on the spot:
//   Coulomb  c_g = tr f_g      -- diagonal lanes + wave shuffle sum (fp64),
//   exchange tr(f_g f_g)       -- transpose through a per-wave LDS tile (fp64),
// so the (naux, o, o) intermediate of meas/rhf.py:45-47 is never formed.
#include "afqmc_hip.h"
#include "common.h"

namespace {


// Walkers w0 .. w0 + WPB - 1 of one batch (nw walkers); NWV waves, each on its
// own g: f_g on an NT x NT grid of 16 x 16 tiles.
template <typename R, int NT, int WPB, int NWV>
__device__ __forceinline__ void energy_body_g(
    int nw, int n, int o, int nspin, int naux, double h0, const double* __restrict__ greens,
    const R* __restrict__ rotp, int ldr, const double* __restrict__ rot_h1,
    double2* __restrict__ eout, int w0, unsigned char* lds_raw) {
    using MF = afq::mfma16<R>;
    using ACC = typename MF::acc;
    using V4 = typename afq::vec4<R>::type;
    constexpr int TC = NT * 16;
    constexpr int XP = TC + 1;
    constexpr int SP = TC + 1;
    constexpr int THREADS = 64 * NWV;
    __shared__ double red[NWV][6 * WPB];
    const int NP = afq::round_up(n, 16);
    R* X = reinterpret_cast<R*>(lds_raw);               // [WPB][2][NP][XP]
    R* S = X + (size_t)WPB * 2 * NP * XP;               // [NWV][TC][SP]
    size_t cs_off = ((size_t)WPB * 2 * NP * XP + NWV * TC * SP) * sizeof(R);
    cs_off = (cs_off + 15) & ~(size_t)15;
    double2* cs = reinterpret_cast<double2*>(lds_raw + cs_off);  // [WPB][naux] (nspin == 2)

    const int nwb = min(WPB, nw - w0);
    const long long ldg = (long long)nspin * o * n;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    R* Sw = S + wave * TC * SP;

    double e1r[WPB], e1i[WPB], exr[WPB], exi[WPB], c2r[WPB], c2i[WPB];
#pragma unroll
    for (int b = 0; b < WPB; ++b) e1r[b] = e1i[b] = exr[b] = exi[b] = c2r[b] = c2i[b] = 0.0;

    for (int s = 0; s < nspin; ++s) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < WPB; ++b) {
            R* Xr = X + (size_t)(b * 2) * NP * XP;
            R* Xi = Xr + (size_t)NP * XP;
            const bool have = b < nwb;
            const double* gre = greens + (long long)(w0 + (have ? b : 0)) * ldg + (long long)s * o * n;
            const double* gim = gre + (long long)nw * ldg;
            for (int e = tid; e < NP * TC; e += THREADS) {
                const int j = e / NP;
                const int p = e - j * NP;
                double vr = 0.0, vi = 0.0;
                if (have && p < n && j < o) {
                    vr = gre[j * n + p];
                    vi = gim[j * n + p];
                }
                Xr[p * XP + j] = (R)vr;
                Xi[p * XP + j] = (R)vi;
            }
            if (have) {
                const double* rh = rot_h1 + (long long)s * o * n;
                for (int e = tid; e < o * n; e += THREADS) {
                    const double h = rh[e];
                    e1r[b] += gre[e] * h;
                    e1i[b] += gim[e] * h;
                }
            }
        }
        __syncthreads();

        const R* Rs = rotp + (long long)s * naux * TC * ldr;
        const int nkb = afq::ceil_div(n, 16);
        for (int g0 = 0; g0 < naux; g0 += NWV) {
            const int g = g0 + wave;
            const bool active = g < naux;
            ACC F[WPB][NT][NT][2];
#pragma unroll
            for (int b = 0; b < WPB; ++b)
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        F[b][ti][tj][0] = MF::zero();
                        F[b][ti][tj][1] = MF::zero();
                    }
            if (active) {
                const R* A = Rs + (long long)g * TC * ldr;
                // A fragments of k-block kb: rows ti * 16 + lr, k = kb * 16 + lk ..
                // + 3 (zero past n); loaded two k-blocks ahead of their use
                auto load_a = [&](int kb, R (&a)[NT][4]) {
                    const int k0 = kb * 16 + lk;
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti) {
                        if (k0 < n) {
                            const V4 v = *reinterpret_cast<const V4*>(A + (long long)(ti * 16 + lr) * ldr + k0);
                            a[ti][0] = v.x; a[ti][1] = v.y; a[ti][2] = v.z; a[ti][3] = v.w;
                        } else {
                            a[ti][0] = a[ti][1] = a[ti][2] = a[ti][3] = R(0);
                        }
                    }
                };
                auto product = [&](int kb, const R (&a)[NT][4]) {
                    const int k0 = kb * 16 + lk;
#pragma unroll
                    for (int b = 0; b < WPB; ++b) {
                        const R* Xr = X + (size_t)(b * 2) * NP * XP;
                        const R* Xi = Xr + (size_t)NP * XP;
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk) {
                            R br[NT], bi[NT];
#pragma unroll
                            for (int tj = 0; tj < NT; ++tj) {
                                br[tj] = Xr[(k0 + kk) * XP + tj * 16 + lr];
                                bi[tj] = Xi[(k0 + kk) * XP + tj * 16 + lr];
                            }
#pragma unroll
                            for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                                for (int tj = 0; tj < NT; ++tj) {
                                    F[b][ti][tj][0] = MF::fma(a[ti][kk], br[tj], F[b][ti][tj][0]);
                                    F[b][ti][tj][1] = MF::fma(a[ti][kk], bi[tj], F[b][ti][tj][1]);
                                }
                        }
                    }
                };
                // three fragment sets in rotation, the loop unrolled by three so
                // that no in-flight set is ever copied
                R a0[NT][4], a1[NT][4], a2[NT][4];
                load_a(0, a0);
                if (nkb > 1) load_a(1, a1);
                auto step = [&](int kb, const R (&cur)[NT][4], R (&ahead)[NT][4]) {
                    if (kb + 2 < nkb) load_a(kb + 2, ahead);
                    product(kb, cur);
                };
                for (int kb = 0; kb < nkb; kb += 3) {
                    step(kb, a0, a2);
                    if (kb + 1 < nkb) step(kb + 1, a1, a0);
                    if (kb + 2 < nkb) step(kb + 2, a2, a1);
                }
            }
#pragma unroll
            for (int b = 0; b < WPB; ++b) {
                // Coulomb: element r of a diagonal tile is on the diagonal when its
                // row MF::row(lane, r) equals its column lane % 16
                double cr = 0.0, ci = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (lr == MF::row(lane, r)) {
                            cr += (double)F[b][t][t][0][r];
                            ci += (double)F[b][t][t][1][r];
                        }
                cr = afq::wave_sum(cr);
                ci = afq::wave_sum(ci);
                if (active && b < nwb) {
                    if (nspin == 1) {
                        c2r[b] += cr * cr - ci * ci;
                        c2i[b] += 2.0 * cr * ci;
                    } else if (s == 0) {
                        if (lane == 0) cs[b * naux + g] = make_double2(cr, ci);
                    } else {
                        const double2 cu = cs[b * naux + g];
                        const double tr = cu.x + cr;
                        const double tim = cu.y + ci;
                        c2r[b] += tr * tr - tim * tim;
                        c2i[b] += 2.0 * tr * tim;
                    }
                }
                // exchange: sum_ij f_ij f_ji, real plane then imaginary plane,
                // through this wave's own LDS tile; a wave's LDS accesses execute
                // in program order, the asm keeps the compiler from moving them
                // across one another (no workgroup barrier in the g loop)
#pragma unroll
                for (int pl = 0; pl < 2; ++pl) {
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = ti * 16 + MF::row(lane, r);
                                const int col = tj * 16 + lr;
                                Sw[row * SP + col] = F[b][ti][tj][pl][r];
                            }
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = ti * 16 + MF::row(lane, r);
                                const int col = tj * 16 + lr;
                                const double t = (double)Sw[col * SP + row];
                                if (pl == 0) {
                                    exr[b] += (double)F[b][ti][tj][0][r] * t;
                                    exi[b] += (double)F[b][ti][tj][1][r] * t;
                                } else {
                                    exr[b] -= (double)F[b][ti][tj][1][r] * t;
                                    exi[b] += (double)F[b][ti][tj][0][r] * t;
                                }
                            }
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                }
            }
        }
    }

#pragma unroll
    for (int b = 0; b < WPB; ++b) {
        const double v[6] = {afq::wave_sum(e1r[b]), afq::wave_sum(e1i[b]), afq::wave_sum(exr[b]),
                             afq::wave_sum(exi[b]), c2r[b], c2i[b]};
        if (lane == 0)
            for (int k = 0; k < 6; ++k) red[wave][6 * b + k] = v[k];
    }
    __syncthreads();
    if (tid < nwb) {
        const int b = tid;
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < NWV; ++q)
            for (int k = 0; k < 6; ++k) v[k] += red[q][6 * b + k];
        double er, ei;
        if (nspin == 1) {
            er = h0 + 2.0 * v[0] + 2.0 * v[4] - v[2];
            ei = 2.0 * v[1] + 2.0 * v[5] - v[3];
        } else {
            er = h0 + v[0] + 0.5 * (v[4] - v[2]);
            ei = v[1] + 0.5 * (v[5] - v[3]);
        }
        eout[w0 + b] = make_double2(er, ei);
    }
}

// Walkers w0 .. w0 + WPB - 1 of one batch (nw walkers); NWV waves.  The
// contraction f_g = R_g G^T runs on stacked rows: a chunk holds q = 16 T / o
// whole g's (o rows each) in T row tiles of 16, so with o not a multiple of 16
// far fewer MFMA rows are padding than with one 16-row tile grid per g; each
// wave takes chunks wave, wave + NWV, ...
template <typename R, int NT, int WPB, int NWV, int T>
__device__ __forceinline__ void energy_body_stacked(
    int nw, int n, int o, int nspin, int naux, double h0, const double* __restrict__ greens,
    const R* __restrict__ rotp, int ldr, const double* __restrict__ rot_h1,
    double2* __restrict__ eout, int w0, unsigned char* lds_raw) {
    using MF = afq::mfma16<R>;
    using ACC = typename MF::acc;
    using V4 = typename afq::vec4<R>::type;
    constexpr int TC = NT * 16;
    constexpr int XP = TC + 1;
    constexpr int SP = TC + 1;
    constexpr int THREADS = 64 * NWV;
    __shared__ double red[NWV][6 * WPB];
    const int NP = afq::round_up(n, 16);
    R* X = reinterpret_cast<R*>(lds_raw);               // [WPB][2][NP][XP]
    R* S = X + (size_t)WPB * 2 * NP * XP;               // [NWV][16 T][SP]
    size_t cs_off = ((size_t)WPB * 2 * NP * XP + NWV * 16 * T * SP) * sizeof(R);
    cs_off = (cs_off + 15) & ~(size_t)15;
    double2* cs = reinterpret_cast<double2*>(lds_raw + cs_off);  // [WPB][naux] (nspin == 2)

    const int nwb = min(WPB, nw - w0);
    const long long ldg = (long long)nspin * o * n;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    R* Sw = S + wave * 16 * T * SP;

    double e1r[WPB], e1i[WPB], exr[WPB], exi[WPB], c2r[WPB], c2i[WPB];
#pragma unroll
    for (int b = 0; b < WPB; ++b) e1r[b] = e1i[b] = exr[b] = exi[b] = c2r[b] = c2i[b] = 0.0;

    for (int s = 0; s < nspin; ++s) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < WPB; ++b) {
            R* Xr = X + (size_t)(b * 2) * NP * XP;
            R* Xi = Xr + (size_t)NP * XP;
            const bool have = b < nwb;
            const double* gre = greens + (long long)(w0 + (have ? b : 0)) * ldg + (long long)s * o * n;
            const double* gim = gre + (long long)nw * ldg;
            for (int e = tid; e < NP * TC; e += THREADS) {
                const int j = e / NP;
                const int p = e - j * NP;
                double vr = 0.0, vi = 0.0;
                if (have && p < n && j < o) {
                    vr = gre[j * n + p];
                    vi = gim[j * n + p];
                }
                Xr[p * XP + j] = (R)vr;
                Xi[p * XP + j] = (R)vi;
            }
            if (have) {
                const double* rh = rot_h1 + (long long)s * o * n;
                for (int e = tid; e < o * n; e += THREADS) {
                    const double h = rh[e];
                    e1r[b] += gre[e] * h;
                    e1i[b] += gim[e] * h;
                }
            }
        }
        __syncthreads();

        const R* Rs = rotp + (long long)s * naux * TC * ldr;
        const int nkb = afq::ceil_div(n, 16);
        // chunks of q whole g's stacked into T row tiles of 16: stacked row
        // r = gl * o + i is row i of f_{g0 + gl}; rows past q * o are zero
        const int q = min(16 * T / o, 16);
        const int nchunk = afq::ceil_div(naux, q);
        // row / o for rows < 64 as (row + 1/2) (1/o) rounded down: the fraction
        // stays at least 1/(2o) away from an integer, far above the rounding
        const float inv_o = 1.0f / (float)o;
        for (int c = wave; c < nchunk; c += NWV) {
            const int g0 = c * q;
            const int qc = min(q, naux - g0);
            ACC F[WPB][T][NT][2];
#pragma unroll
            for (int b = 0; b < WPB; ++b)
#pragma unroll
                for (int t = 0; t < T; ++t)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) {
                        F[b][t][tj][0] = MF::zero();
                        F[b][t][tj][1] = MF::zero();
                    }
            // this lane's A rows (t * 16 + lr of the stack) in the padded
            // (g, TC, ldr) layout of the rotated Cholesky vectors
            const R* arow[T];
            bool aok[T];
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const int r = t * 16 + lr;
                const int gl = (int)(((float)r + 0.5f) * inv_o);
                aok[t] = gl < qc;
                arow[t] = Rs + ((long long)(g0 + (aok[t] ? gl : 0)) * TC + (aok[t] ? r - gl * o : 0)) * ldr;
            }
            // A fragments of k-block kb: k = kb * 16 + lk .. + 3 (zero past n);
            // loaded two k-blocks ahead of their use
            auto load_a = [&](int kb, R (&a)[T][4]) {
                const int k0 = kb * 16 + lk;
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    if (aok[t] && k0 < n) {
                        const V4 v = *reinterpret_cast<const V4*>(arow[t] + k0);
                        a[t][0] = v.x; a[t][1] = v.y; a[t][2] = v.z; a[t][3] = v.w;
                    } else {
                        a[t][0] = a[t][1] = a[t][2] = a[t][3] = R(0);
                    }
                }
            };
            auto product = [&](int kb, const R (&a)[T][4]) {
                const int k0 = kb * 16 + lk;
#pragma unroll
                for (int b = 0; b < WPB; ++b) {
                    const R* Xr = X + (size_t)(b * 2) * NP * XP;
                    const R* Xi = Xr + (size_t)NP * XP;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        R br[NT], bi[NT];
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj) {
                            br[tj] = Xr[(k0 + kk) * XP + tj * 16 + lr];
                            bi[tj] = Xi[(k0 + kk) * XP + tj * 16 + lr];
                        }
#pragma unroll
                        for (int t = 0; t < T; ++t)
#pragma unroll
                            for (int tj = 0; tj < NT; ++tj) {
                                F[b][t][tj][0] = MF::fma(a[t][kk], br[tj], F[b][t][tj][0]);
                                F[b][t][tj][1] = MF::fma(a[t][kk], bi[tj], F[b][t][tj][1]);
                            }
                    }
                }
            };
            // fragment sets in rotation (three, two k-blocks ahead; two, one
            // ahead, for four row tiles, whose accumulators need the
            // registers), the loop unrolled by the set count so that no
            // in-flight set is ever copied
            if constexpr (T == 4) {
                R a0[T][4], a1[T][4];
                load_a(0, a0);
                for (int kb = 0; kb < nkb; kb += 2) {
                    if (kb + 1 < nkb) load_a(kb + 1, a1);
                    product(kb, a0);
                    if (kb + 1 < nkb) {
                        if (kb + 2 < nkb) load_a(kb + 2, a0);
                        product(kb + 1, a1);
                    }
                }
            } else {
                R a0[T][4], a1[T][4], a2[T][4];
                load_a(0, a0);
                if (nkb > 1) load_a(1, a1);
                auto step = [&](int kb, const R (&cur)[T][4], R (&ahead)[T][4]) {
                    if (kb + 2 < nkb) load_a(kb + 2, ahead);
                    product(kb, cur);
                };
                for (int kb = 0; kb < nkb; kb += 3) {
                    step(kb, a0, a2);
                    if (kb + 1 < nkb) step(kb + 1, a1, a0);
                    if (kb + 2 < nkb) step(kb + 2, a2, a1);
                }
            }
#pragma unroll
            for (int b = 0; b < WPB; ++b) {
                // Each plane (real, then imaginary) goes through this wave's own
                // LDS tile (T * 16 rows): lane gl < qc sums the diagonal of
                // f_{g0 + gl} (Coulomb), and every lane adds f_ij f_ji of its
                // entries (exchange).  A wave's LDS accesses execute in program
                // order; the asm keeps the compiler from moving them across one
                // another (no workgroup barrier in the chunk loop).
                double cre = 0.0, cim = 0.0;
                // the lane id through an empty asm: the entry and partner
                // indices below are recomputed per chunk instead of being
                // kept in registers across the chunk loop
                int lx = lane;
                asm volatile("" : "+v"(lx));
                const int lrx = lx & 15;
#pragma unroll
                for (int pl = 0; pl < 2; ++pl) {
#pragma unroll
                    for (int t = 0; t < T; ++t)
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int row = t * 16 + MF::row(lx, r);
                                const int col = tj * 16 + lrx;
                                Sw[row * SP + col] = F[b][t][tj][pl][r];
                            }
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                    {
                        // lane l holds the diagonal entry of stacked row l; the
                        // trace of f_{g0 + gl} is the sum over its o rows
                        const int glx = (int)(((float)lx + 0.5f) * inv_o);
                        const double d = lx < qc * o ? (double)Sw[lx * SP + (lx - glx * o)] : 0.0;
                        for (int gl = 0; gl < qc; ++gl) {
                            const double sg = afq::wave_sum(glx == gl ? d : 0.0);
                            if (lx == gl) {
                                if (pl == 0)
                                    cre = sg;
                                else
                                    cim = sg;
                            }
                        }
                    }
#pragma unroll
                    for (int t = 0; t < T; ++t)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = t * 16 + MF::row(lx, r);
                            const int gl = (int)(((float)row + 0.5f) * inv_o);
                            const int i = row - gl * o;
#pragma unroll
                            for (int tj = 0; tj < NT; ++tj) {
                                const int col = tj * 16 + lrx;
                                if (gl < qc && col < o) {
                                    const double x = (double)Sw[(gl * o + col) * SP + i];
                                    // converted at each use (an empty asm per plane keeps
                                    // the compiler from holding both planes' conversions)
                                    R fr = F[b][t][tj][0][r], fi = F[b][t][tj][1][r];
                                    asm volatile("" : "+v"(fr), "+v"(fi));
                                    if (pl == 0) {
                                        exr[b] += (double)fr * x;
                                        exi[b] += (double)fi * x;
                                    } else {
                                        exr[b] -= (double)fi * x;
                                        exi[b] += (double)fr * x;
                                    }
                                }
                            }
                        }
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                }
                // Coulomb (c_g = tr f_g; per lane, summed over lanes at the end)
                if (lane < qc && b < nwb) {
                    const int g = g0 + lane;
                    if (nspin == 1) {
                        c2r[b] += cre * cre - cim * cim;
                        c2i[b] += 2.0 * cre * cim;
                    } else if (s == 0) {
                        cs[b * naux + g] = make_double2(cre, cim);
                    } else {
                        const double2 cu = cs[b * naux + g];
                        const double tr = cu.x + cre;
                        const double tim = cu.y + cim;
                        c2r[b] += tr * tr - tim * tim;
                        c2i[b] += 2.0 * tr * tim;
                    }
                }
            }
        }
    }

#pragma unroll
    for (int b = 0; b < WPB; ++b) {
        const double v[6] = {afq::wave_sum(e1r[b]), afq::wave_sum(e1i[b]), afq::wave_sum(exr[b]),
                             afq::wave_sum(exi[b]), afq::wave_sum(c2r[b]), afq::wave_sum(c2i[b])};
        if (lane == 0)
            for (int k = 0; k < 6; ++k) red[wave][6 * b + k] = v[k];
    }
    __syncthreads();
    if (tid < nwb) {
        const int b = tid;
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < NWV; ++q)
            for (int k = 0; k < 6; ++k) v[k] += red[q][6 * b + k];
        double er, ei;
        if (nspin == 1) {
            er = h0 + 2.0 * v[0] + 2.0 * v[4] - v[2];
            ei = 2.0 * v[1] + 2.0 * v[5] - v[3];
        } else {
            er = h0 + v[0] + 0.5 * (v[4] - v[2]);
            ei = v[1] + 0.5 * (v[5] - v[3]);
        }
        eout[w0 + b] = make_double2(er, ei);
    }
}

// Kernel shape of the f_g contraction for o occupied orbitals: 4 (stacked
// chunks of four row tiles) when 64 stacked rows hold more than twice the
// whole g's of 32 (o = 7, 9, 11, 12, 17..21: the per-g tile grid would be
// mostly padding), else 0 (one NT x NT tile grid per g, which measures faster
// when stacking gains little).  A function of o alone, so a system gets the
// same partition of g's (and the same sums) in the single-system and the
// grouped kernels.
__host__ __device__ inline int stack_tiles(int o) {
    const int q2 = min(32 / o, 16), q4 = min(64 / o, 16);
    return q4 > 2 * q2 ? 4 : 0;
}

template <typename R, int NT, int WPB, int NWV, int T>
__global__ __launch_bounds__(64 * NWV, 1) void energy_kernel(
    int nw, int n, int o, int nspin, int naux, double h0, const double* __restrict__ greens,
    const R* __restrict__ rotp, int ldr, const double* __restrict__ rot_h1,
    double2* __restrict__ eout) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    if constexpr (T == 0)
        energy_body_g<R, NT, WPB, NWV>(nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, eout, blockIdx.x * WPB,
                                       lds_raw);
    else
        energy_body_stacked<R, NT, WPB, NWV, T>(nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, eout,
                                                blockIdx.x * WPB, lds_raw);
}

// One workgroup per walker of every system (restricted walkers, fp32
// contractions); every system has ceil(o / 16) == NT.  Eight waves and the
// chunk shape stack_tiles(o), as the single-system fp32 kernel has (same g
// partition, so the same sums); the LDS is sized for four row tiles.
constexpr int F32_WAVES = 8;
template <int NT>
__global__ __launch_bounds__(64 * F32_WAVES, 1) void energy_f32_grouped_kernel(
    int count, const afq_system_step* __restrict__ S) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    if (stack_tiles(d.o) == 4)
        energy_body_stacked<float, NT, 1, F32_WAVES, 4>(d.nw, d.n, d.o, 1, d.naux, d.h0, d.greens, d.rot_chol,
                                                        d.ldr, d.rot_h1, (double2*)d.energy, b - d.walker0, lds_raw);
    else
        energy_body_g<float, NT, 1, F32_WAVES>(d.nw, d.n, d.o, 1, d.naux, d.h0, d.greens, d.rot_chol, d.ldr,
                                               d.rot_h1, (double2*)d.energy, b - d.walker0, lds_raw);
}

template <typename R, int NT, int WPB, int NWV, int T>
size_t energy_lds(int n, int nspin, int naux) {
    constexpr int TC = NT * 16;
    const int NP = afq::round_up(n, 16);
    const int srows = T == 0 ? TC : 16 * T;  // rows of a wave's exchange tile
    size_t b = sizeof(R) * ((size_t)WPB * 2 * NP * (TC + 1) + (size_t)NWV * srows * (TC + 1));
    b = (b + 15) & ~(size_t)15;
    if (nspin == 2) b += sizeof(double2) * (size_t)WPB * naux;
    return b;
}

template <typename R, int NT, int WPB, int NWV, int T>
bool energy_fits(int n, int nspin, int naux) {
    return energy_lds<R, NT, WPB, NWV, T>(n, nspin, naux) + NWV * 6 * WPB * sizeof(double) <= 160 * 1024;
}

template <typename R, int NT, int WPB, int NWV, int T>
int launch_energy(const char* name, int nw, int n, int o, int nspin, int naux, double h0,
                  const double* greens, const R* rotp, int ldr, const double* rot_h1,
                  double* energy, hipStream_t st) {
    const size_t lds = energy_lds<R, NT, WPB, NWV, T>(n, nspin, naux);
    AFQ_REQUIRE((energy_fits<R, NT, WPB, NWV, T>(n, nspin, naux)),
                "%s: n=%d o=%d naux=%d needs more than 160 KB of LDS", name, n, o, naux);
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)energy_kernel<R, NT, WPB, NWV, T>);
    hipLaunchKernelGGL((energy_kernel<R, NT, WPB, NWV, T>), dim3((unsigned)afq::ceil_div(nw, WPB)),
                       dim3(64 * NWV), lds, st, nw, n, o, nspin, naux, h0, greens, rotp, ldr,
                       rot_h1, (double2*)energy);
    return afq::check_launch(name);
}

// T = 0: per-g tile grid, T = 4: stacked chunks.  fp32: eight waves; two
// walkers per workgroup when they fit in LDS (halves the rotated-Cholesky
// traffic) and registers (not stacked with two column tiles), else one (the
// walker count changes no sum).  fp64: one walker, eight waves when they fit
// (per-g grid only), else four.
template <typename R, int NT, int T>
int energy_dispatch(const char* name, int nw, int n, int o, int nspin, int naux, double h0,
                    const double* greens, const R* rotp, int ldr, const double* rot_h1, double* energy,
                    hipStream_t st) {
    if constexpr (sizeof(R) == sizeof(float)) {
        constexpr int W = F32_WAVES;
        if constexpr (T == 0 || NT == 1) {
            if (energy_fits<R, NT, 2, W, T>(n, nspin, naux))
                return launch_energy<R, NT, 2, W, T>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy,
                                                     st);
        }
        return launch_energy<R, NT, 1, W, T>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy, st);
    } else {
        if constexpr (T == 0) {
            if (energy_fits<R, NT, 1, 8, T>(n, nspin, naux))
                return launch_energy<R, NT, 1, 8, T>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy,
                                                     st);
        }
        return launch_energy<R, NT, 1, 4, T>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy, st);
    }
}

template <typename R>
int energy_entry(const char* name, int nw, int n, int o, int nspin, int naux, double h0,
                 const double* greens, const R* rotp, int ldr, const double* rot_h1,
                 double* energy, hipStream_t st) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n && naux > 0, "%s: bad sizes", name);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "%s: nspin must be 1 or 2", name);
    AFQ_REQUIRE(ldr >= n && ldr % 4 == 0, "%s: ldr=%d must be >= n and a multiple of 4", name, ldr);
    AFQ_REQUIRE(greens && rotp && rot_h1 && energy, "%s: null pointer", name);
    AFQ_REQUIRE(o <= 32, "%s: o=%d exceeds the supported maximum of 32", name, o);
    // stacked chunks where they pack (fp32 always, fp64 with two column tiles:
    // with one, its four-wave workgroups measure slower than the per-g grid)
    // and their LDS tile fits; the per-g tile grid otherwise
    constexpr bool f32 = sizeof(R) == sizeof(float);
    const bool stacked = stack_tiles(o) == 4;
    if (o <= 16) {
        if constexpr (f32) {
            if (stacked && energy_fits<R, 1, 1, F32_WAVES, 4>(n, nspin, naux))
                return energy_dispatch<R, 1, 4>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy, st);
        }
        return energy_dispatch<R, 1, 0>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy, st);
    }
    if (stacked && energy_fits<R, 2, 1, f32 ? F32_WAVES : 4, 4>(n, nspin, naux))
        return energy_dispatch<R, 2, 4>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy, st);
    return energy_dispatch<R, 2, 0>(name, nw, n, o, nspin, naux, h0, greens, rotp, ldr, rot_h1, energy, st);
}

// ---------------------------------------------------------------------------
// Split-bf16 variant: f_g = R_g G^T with both operands carried as bf16
// hi + lo and three v_mfma_f32_16x16x32_bf16 per k-block (fp32 sums within a
// g, fp64 sums over g).  G of WPB walkers sits in LDS as four bf16 planes
// (re hi/lo, im hi/lo) in its stored order (occupied index j, then orbital p),
// which is the B-fragment order (8 consecutive p of one column j per lane).
// R_g (hi/lo, rows padded to TC, pitch ldr a multiple of 32) is read from
// L2 as A fragments, two k-blocks ahead.  Eight waves, each on its own g; the
// exchange transpose goes through a per-wave LDS tile, ordered within the wave
// (no workgroup barrier inside the g loop).
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int THREADS_B = 512;
constexpr int KMAX_E = 8;  // k-blocks of 32 orbitals: n <= 256

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int NT, int WPB>
size_t energy_bf16x3_lds(int n, int nspin, int naux) {
    constexpr int TC = NT * 16;
    const int pp = afq::round_up(n, 32) + 8;
    size_t b = (size_t)WPB * 4 * TC * pp * 2 + (size_t)8 * TC * (TC + 1) * 4;
    b = (b + 15) & ~(size_t)15;
    if (nspin == 2) b += sizeof(double2) * (size_t)WPB * naux;
    return b;
}

// Walkers w0 .. w0 + WPB - 1 of one batch (nw walkers), in a workgroup of
// THREADS_B threads whose dynamic LDS (lds_raw) holds energy_bf16x3_lds for
// this ldr or a larger one.
template <int NT, int WPB>
__device__ __forceinline__ void energy_bf16x3_body(
    int nw, int n, int o, int nspin, int naux, double h0, const double* __restrict__ greens,
    const uint16_t* __restrict__ rot_hi, const uint16_t* __restrict__ rot_lo, int ldr,
    const double* __restrict__ rot_h1, double2* __restrict__ eout, int w0, unsigned char* lds_raw) {
    using MF = afq::mfma16<float>;
    constexpr int TC = NT * 16;
    constexpr int SP = TC + 1;
    __shared__ double red[8][6 * WPB];
    const int nk = ldr / 32;
    const int PP = ldr + 8;  // bf16 pitch of one column of G
    __bf16* X = reinterpret_cast<__bf16*>(lds_raw);               // [WPB][4][TC][PP]
    float* S = reinterpret_cast<float*>(lds_raw + (size_t)WPB * 4 * TC * PP * 2);  // [8][TC][SP]
    size_t cs_off = (size_t)WPB * 4 * TC * PP * 2 + (size_t)8 * TC * SP * 4;
    cs_off = (cs_off + 15) & ~(size_t)15;
    double2* cs = reinterpret_cast<double2*>(lds_raw + cs_off);   // [WPB][naux] (nspin == 2)

    const int nwb = min(WPB, nw - w0);
    const long long ldg = (long long)nspin * o * n;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = lane & 15;
    const int q = lane >> 4;
    float* Sw = S + wave * TC * SP;

    double e1r[WPB], e1i[WPB], exr[WPB], exi[WPB], c2r[WPB], c2i[WPB];
#pragma unroll
    for (int b = 0; b < WPB; ++b) e1r[b] = e1i[b] = exr[b] = exi[b] = c2r[b] = c2i[b] = 0.0;

    for (int s = 0; s < nspin; ++s) {
        __syncthreads();
        // G of each walker -> four bf16 planes; zero for j >= o and p >= n
#pragma unroll
        for (int b = 0; b < WPB; ++b) {
            const bool have = b < nwb;
            const double* gre = greens + (long long)(w0 + (have ? b : 0)) * ldg + (long long)s * o * n;
            const double* gim = gre + (long long)nw * ldg;
            __bf16* Xb = X + (size_t)b * 4 * TC * PP;
            for (int e = tid; e < TC * ldr; e += THREADS_B) {
                const int j = e / ldr;
                const int p = e - j * ldr;
                const bool ok = have && j < o && p < n;
                const int idx = ok ? j * n + p : 0;
                const double vr = ok ? gre[idx] : 0.0;
                const double vi = ok ? gim[idx] : 0.0;
                const float fr = (float)vr, fi = (float)vi;
                const __bf16 hr = (__bf16)fr, hi = (__bf16)fi;
                Xb[(0 * TC + j) * PP + p] = hr;
                Xb[(1 * TC + j) * PP + p] = (__bf16)(fr - (float)hr);
                Xb[(2 * TC + j) * PP + p] = hi;
                Xb[(3 * TC + j) * PP + p] = (__bf16)(fi - (float)hi);
            }
            if (have) {
                const double* rh = rot_h1 + (long long)s * o * n;
                for (int e = tid; e < o * n; e += THREADS_B) {
                    const double h = rh[e];
                    e1r[b] += gre[e] * h;
                    e1i[b] += gim[e] * h;
                }
            }
        }
        __syncthreads();

        const long long rs = (long long)s * naux * TC * ldr;
        auto read_b = [&](int b, int plane, int tj, int kb) {
            return *reinterpret_cast<const bf16x8*>(X + ((size_t)(b * 4 + plane) * TC + tj * 16 + lr) * PP +
                                                   kb * 32 + 8 * q);
        };
        for (int g = wave; g < naux; g += 8) {
            const uint16_t* Ah = rot_hi + rs + (long long)g * TC * ldr;
            const uint16_t* Al = rot_lo + rs + (long long)g * TC * ldr;
            f32x4 F[WPB][NT][NT][2];
#pragma unroll
            for (int b = 0; b < WPB; ++b)
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj) F[b][ti][tj][0] = F[b][ti][tj][1] = MF::zero();
            // A fragments of R_g (rows ti*16 + lr, k = kb*32 + 8q..), two k-blocks ahead
            bf16x8 ring[3][NT][2];
            auto issue = [&](int kb) {
#pragma unroll
                for (int ti = 0; ti < NT; ++ti) {
                    const long long off = (long long)(ti * 16 + lr) * ldr + (kb < nk ? kb : 0) * 32 + 8 * q;
                    ring[kb % 3][ti][0] = *reinterpret_cast<const bf16x8*>(Ah + off);
                    ring[kb % 3][ti][1] = *reinterpret_cast<const bf16x8*>(Al + off);
                }
            };
            issue(0);
            if (nk > 1) issue(1);
            // fully unrolled to KMAX_E k-blocks so that every ring index is a constant
#pragma unroll
            for (int kb = 0; kb < KMAX_E; ++kb) {
                if (kb >= nk) break;
                if (kb + 2 < nk) issue(kb + 2);
                bf16x8 ah[NT], al[NT];
#pragma unroll
                for (int ti = 0; ti < NT; ++ti) {
                    ah[ti] = ring[kb % 3][ti][0];
                    al[ti] = ring[kb % 3][ti][1];
                }
#pragma unroll
                for (int b = 0; b < WPB; ++b)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            const bf16x8 bh = read_b(b, 2 * c, tj, kb);
                            const bf16x8 bl = read_b(b, 2 * c + 1, tj, kb);
#pragma unroll
                            for (int ti = 0; ti < NT; ++ti) {
                                f32x4& acc = F[b][ti][tj][c];
                                acc = mfma_bf16(al[ti], bh, acc);
                                acc = mfma_bf16(ah[ti], bl, acc);
                                acc = mfma_bf16(ah[ti], bh, acc);
                            }
                        }
                // keeps the next k-block's fragment reads out of this one
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int b = 0; b < WPB; ++b) {
                // Coulomb: element r of a diagonal tile is on the diagonal when
                // its row MF::row(lane, r) equals its column lane % 16
                double cr = 0.0, ci = 0.0;
#pragma unroll
                for (int t = 0; t < NT; ++t)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (lr == MF::row(lane, r)) {
                            cr += (double)F[b][t][t][0][r];
                            ci += (double)F[b][t][t][1][r];
                        }
                cr = afq::wave_sum(cr);
                ci = afq::wave_sum(ci);
                if (b < nwb) {
                    if (nspin == 1) {
                        c2r[b] += cr * cr - ci * ci;
                        c2i[b] += 2.0 * cr * ci;
                    } else if (s == 0) {
                        if (lane == 0) cs[b * naux + g] = make_double2(cr, ci);
                    } else {
                        const double2 cu = cs[b * naux + g];
                        const double tr = cu.x + cr;
                        const double tim = cu.y + ci;
                        c2r[b] += tr * tr - tim * tim;
                        c2i[b] += 2.0 * tr * tim;
                    }
                }
                // exchange sum_ij f_ij f_ji through this wave's LDS tile; the
                // wave's LDS accesses execute in program order, the asm keeps
                // the compiler from moving them across one another
#pragma unroll
                for (int pl = 0; pl < 2; ++pl) {
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                            for (int r = 0; r < 4; ++r)
                                Sw[(ti * 16 + MF::row(lane, r)) * SP + tj * 16 + lr] = F[b][ti][tj][pl][r];
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const double t = (double)Sw[(tj * 16 + lr) * SP + ti * 16 + MF::row(lane, r)];
                                if (pl == 0) {
                                    exr[b] += (double)F[b][ti][tj][0][r] * t;
                                    exi[b] += (double)F[b][ti][tj][1][r] * t;
                                } else {
                                    exr[b] -= (double)F[b][ti][tj][1][r] * t;
                                    exi[b] += (double)F[b][ti][tj][0][r] * t;
                                }
                            }
                    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
                }
            }
        }
    }

#pragma unroll
    for (int b = 0; b < WPB; ++b) {
        const double v[6] = {afq::wave_sum(e1r[b]), afq::wave_sum(e1i[b]), afq::wave_sum(exr[b]),
                             afq::wave_sum(exi[b]), c2r[b], c2i[b]};
        if (lane == 0)
            for (int k = 0; k < 6; ++k) red[wave][6 * b + k] = v[k];
    }
    __syncthreads();
    if (tid < nwb) {
        const int b = tid;
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int w8 = 0; w8 < 8; ++w8)
            for (int k = 0; k < 6; ++k) v[k] += red[w8][6 * b + k];
        double er, ei;
        if (nspin == 1) {
            er = h0 + 2.0 * v[0] + 2.0 * v[4] - v[2];
            ei = 2.0 * v[1] + 2.0 * v[5] - v[3];
        } else {
            er = h0 + v[0] + 0.5 * (v[4] - v[2]);
            ei = v[1] + 0.5 * (v[5] - v[3]);
        }
        eout[w0 + b] = make_double2(er, ei);
    }
}

template <int NT, int WPB>
__global__ __launch_bounds__(THREADS_B, 1) void energy_bf16x3_kernel(
    int nw, int n, int o, int nspin, int naux, double h0, const double* __restrict__ greens,
    const uint16_t* __restrict__ rot_hi, const uint16_t* __restrict__ rot_lo, int ldr,
    const double* __restrict__ rot_h1, double2* __restrict__ eout) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    energy_bf16x3_body<NT, WPB>(nw, n, o, nspin, naux, h0, greens, rot_hi, rot_lo, ldr, rot_h1, eout,
                                blockIdx.x * WPB, lds_raw);
}

// Every system of a bucket of equal ceil(o / 16) == NT (restricted walkers):
// workgroup -> (system, WPB walkers) by the prefix sums tblock0 of
// ceil(nw / WPB); the system's rot_chol points at its hi plane, followed by
// the lo plane, (naux * 16 NT, ldr) each, ldr = n rounded up to 32.  The
// same body and walker grouping as the single-system kernel with this WPB.
template <int NT, int WPB>
__global__ __launch_bounds__(THREADS_B, 1) void energy_bf16x3_grouped_kernel(int count,
                                                                             const afq_system_step* __restrict__ S) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].tblock0; });
    const afq_system_step& d = S[si];
    const uint16_t* rh = reinterpret_cast<const uint16_t*>(d.rot_chol);
    const uint16_t* rl = rh + (long long)d.naux * (16 * NT) * d.ldr;
    energy_bf16x3_body<NT, WPB>(d.nw, d.n, d.o, 1, d.naux, d.h0, d.greens, rh, rl, d.ldr, d.rot_h1,
                                (double2*)d.energy, (b - d.tblock0) * WPB, lds_raw);
}

template <int NT, int WPB>
int launch_energy_bf16x3(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
                         const uint16_t* rh, const uint16_t* rl, int ldr, const double* rot_h1,
                         double* energy, hipStream_t st) {
    const size_t lds = energy_bf16x3_lds<NT, WPB>(ldr, nspin, naux);
    AFQ_REQUIRE(lds + 8 * 6 * WPB * sizeof(double) <= 160 * 1024,
                "afq_energy_bf16x3: n=%d o=%d naux=%d needs more than 160 KB of LDS", n, o, naux);
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)energy_bf16x3_kernel<NT, WPB>);
    hipLaunchKernelGGL((energy_bf16x3_kernel<NT, WPB>), dim3((unsigned)afq::ceil_div(nw, WPB)),
                       dim3(THREADS_B), lds, st, nw, n, o, nspin, naux, h0, greens, rh, rl, ldr, rot_h1,
                       (double2*)energy);
    return afq::check_launch("afq_energy_bf16x3");
}

}  // namespace

extern "C" int afq_energy(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
                          const double* rot_chol_pad, int ldr, const double* rot_h1, double* energy,
                          void* stream) {
    return energy_entry<double>("afq_energy", nw, n, o, nspin, naux, h0, greens, rot_chol_pad, ldr,
                                rot_h1, energy, (hipStream_t)stream);
}

extern "C" int afq_energy_f32(int nw, int n, int o, int nspin, int naux, double h0,
                              const double* greens, const float* rot_chol_pad, int ldr,
                              const double* rot_h1, double* energy, void* stream) {
    return energy_entry<float>("afq_energy_f32", nw, n, o, nspin, naux, h0, greens, rot_chol_pad,
                               ldr, rot_h1, energy, (hipStream_t)stream);
}

extern "C" int afq_energy_bf16x3(int nw, int n, int o, int nspin, int naux, double h0,
                                 const double* greens, const unsigned short* rot_hi,
                                 const unsigned short* rot_lo, int ldr, const double* rot_h1,
                                 double* energy, void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n && naux > 0, "afq_energy_bf16x3: bad sizes");
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "afq_energy_bf16x3: nspin must be 1 or 2");
    AFQ_REQUIRE(ldr >= n && ldr % 32 == 0, "afq_energy_bf16x3: ldr=%d must be >= n and a multiple of 32", ldr);
    AFQ_REQUIRE(ldr <= 32 * KMAX_E, "afq_energy_bf16x3: n=%d exceeds the supported maximum of %d", n, 32 * KMAX_E);
    AFQ_REQUIRE(greens && rot_hi && rot_lo && rot_h1 && energy, "afq_energy_bf16x3: null pointer");
    AFQ_REQUIRE(o <= 32, "afq_energy_bf16x3: o=%d exceeds the supported maximum of 32", o);
    hipStream_t st = (hipStream_t)stream;
    const uint16_t* rh = (const uint16_t*)rot_hi;
    const uint16_t* rl = (const uint16_t*)rot_lo;
    // two walkers per workgroup (halves the R_g traffic) when they fit in LDS
    // (the per-g Coulomb store of the unrestricted form takes room too)
    auto fits2 = [&](size_t lds) { return lds + 8 * 6 * 2 * sizeof(double) <= 160 * 1024; };
    if (o <= 16) {
        if (fits2(energy_bf16x3_lds<1, 2>(ldr, nspin, naux)))
            return launch_energy_bf16x3<1, 2>(nw, n, o, nspin, naux, h0, greens, rh, rl, ldr, rot_h1, energy, st);
        return launch_energy_bf16x3<1, 1>(nw, n, o, nspin, naux, h0, greens, rh, rl, ldr, rot_h1, energy, st);
    }
    if (fits2(energy_bf16x3_lds<2, 2>(ldr, nspin, naux)))
        return launch_energy_bf16x3<2, 2>(nw, n, o, nspin, naux, h0, greens, rh, rl, ldr, rot_h1, energy, st);
    return launch_energy_bf16x3<2, 1>(nw, n, o, nspin, naux, h0, greens, rh, rl, ldr, rot_h1, energy, st);
}

extern "C" int afq_energy_f32_grouped(int count, const afq_system_step* sys, int total_walkers,
                                      int max_n, int nt, void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_energy_f32_grouped: count=%d total_walkers=%d",
                count, total_walkers);
    AFQ_REQUIRE(sys, "afq_energy_f32_grouped: null pointer");
    AFQ_REQUIRE(nt == 1 || nt == 2, "afq_energy_f32_grouped: nt=%d must be 1 or 2", nt);
    AFQ_REQUIRE(max_n > 0 && max_n <= 256, "afq_energy_f32_grouped: max_n=%d", max_n);
    hipStream_t st = (hipStream_t)stream;
    if (nt == 1) {
        const size_t lds = energy_lds<float, 1, 1, F32_WAVES, 4>(max_n, 1, 0);
        AFQ_REQUIRE((energy_fits<float, 1, 1, F32_WAVES, 4>(max_n, 1, 0)), "afq_energy_f32_grouped: max_n=%d too large",
                    max_n);
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)energy_f32_grouped_kernel<1>);
        hipLaunchKernelGGL(energy_f32_grouped_kernel<1>, dim3((unsigned)total_walkers), dim3(64 * F32_WAVES), lds, st,
                           count, sys);
    } else {
        const size_t lds = energy_lds<float, 2, 1, F32_WAVES, 4>(max_n, 1, 0);
        AFQ_REQUIRE((energy_fits<float, 2, 1, F32_WAVES, 4>(max_n, 1, 0)), "afq_energy_f32_grouped: max_n=%d too large",
                    max_n);
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)energy_f32_grouped_kernel<2>);
        hipLaunchKernelGGL(energy_f32_grouped_kernel<2>, dim3((unsigned)total_walkers), dim3(64 * F32_WAVES), lds, st,
                           count, sys);
    }
    return afq::check_launch("afq_energy_f32_grouped");
}

extern "C" int afq_energy_bf16x3_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n, int nt,
                                         int wpb, void* stream) {
    AFQ_REQUIRE(count > 0 && total_blocks > 0, "afq_energy_bf16x3_grouped: count=%d total_blocks=%d", count,
                total_blocks);
    AFQ_REQUIRE(sys, "afq_energy_bf16x3_grouped: null pointer");
    AFQ_REQUIRE(nt == 1 || nt == 2, "afq_energy_bf16x3_grouped: nt=%d must be 1 or 2", nt);
    AFQ_REQUIRE(wpb == 1 || wpb == 2, "afq_energy_bf16x3_grouped: wpb=%d must be 1 or 2", wpb);
    AFQ_REQUIRE(max_n > 0 && max_n <= 32 * KMAX_E, "afq_energy_bf16x3_grouped: max_n=%d not in 1..%d", max_n,
                32 * KMAX_E);
    const int ldr = afq::round_up(max_n, 32);
    hipStream_t st = (hipStream_t)stream;
#define AFQ_ENERGY_GROUPED(NT_, WPB_)                                                                          \
    if (nt == NT_ && wpb == WPB_) {                                                                            \
        const size_t lds = energy_bf16x3_lds<NT_, WPB_>(ldr, 1, 0);                                            \
        AFQ_REQUIRE(lds + 8 * 6 * WPB_ * sizeof(double) <= 160 * 1024,                                        \
                    "afq_energy_bf16x3_grouped: max_n=%d nt=%d wpb=%d needs more than 160 KB of LDS", max_n, nt, \
                    wpb);                                                                                      \
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)energy_bf16x3_grouped_kernel<NT_, WPB_>);        \
        hipLaunchKernelGGL((energy_bf16x3_grouped_kernel<NT_, WPB_>), dim3((unsigned)total_blocks),            \
                           dim3(THREADS_B), lds, st, count, sys);                                              \
        return afq::check_launch("afq_energy_bf16x3_grouped");                                                 \
    }
    AFQ_ENERGY_GROUPED(1, 1)
    AFQ_ENERGY_GROUPED(1, 2)
    AFQ_ENERGY_GROUPED(2, 1)
    AFQ_ENERGY_GROUPED(2, 2)
#undef AFQ_ENERGY_GROUPED
    afq::set_error("afq_energy_bf16x3_grouped: no kernel for nt=%d wpb=%d", nt, wpb);
    return -1;
}

extern "C" int afq_energy_bf16x3_grouped_fits(int max_n, int nt, int wpb) {
    const int ldr = afq::round_up(max_n, 32);
    size_t lds = 0;
    if (nt == 1)
        lds = wpb == 1 ? energy_bf16x3_lds<1, 1>(ldr, 1, 0) : energy_bf16x3_lds<1, 2>(ldr, 1, 0);
    else
        lds = wpb == 1 ? energy_bf16x3_lds<2, 1>(ldr, 1, 0) : energy_bf16x3_lds<2, 2>(ldr, 1, 0);
    return lds + 8 * 6 * wpb * sizeof(double) <= 160 * 1024 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Local energies from the exchange as a quadratic form (meas/rhf.py:36-50,
// 75-105 with the sum over g regrouped): per walker
//   e1   = sum_s tr(rot_h1_s G_s)                      (one-body)
//   coul = sum_g c_g^2, c = c_scale x the force-bias contraction (c summed over
//          spin sectors: sum_g (c_up + c_dn)_g^2 for unrestricted walkers)
//   exc  = sum_s vec(G_s)^T Z_s, Z_s = G_s Y_s the caller's GEMM of the sector's
//          rows with Y_s[(a,q),(b,p)] = sum_g R[g,a,p] R[g,b,q]   (no conjugation)
//   E    = h0 + 2 e1 + 2 coul - exc (restricted), h0 + e1 + (coul - exc) / 2
// One workgroup per walker, fp64 sums: each thread strides its share in a fixed
// order, then a fixed tree over the workgroup (the result does not depend on
// the launch).
namespace {

constexpr int THREADS_A = 256;

__device__ void energy_assemble_walker(int w, int nw, int on, int nspin, int naux, double h0, double c_scale,
                                       const double* __restrict__ G, const double* __restrict__ Z,
                                       const double* __restrict__ c, const double* __restrict__ h1,
                                       double2* __restrict__ E, double (*red)[THREADS_A]) {
    const int tid = threadIdx.x;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // e1 re/im, exc re/im, coul re/im
    const long long rowg = (long long)nspin * on;
    for (int s = 0; s < nspin; ++s) {
        const double* gr = G + (long long)w * rowg + (long long)s * on;
        const double* gi = G + (long long)(nw + w) * rowg + (long long)s * on;
        const double* zr = Z + (long long)s * 2 * nw * on + (long long)w * on;
        const double* zi = Z + (long long)s * 2 * nw * on + (long long)(nw + w) * on;
        const double* hs = h1 + (long long)s * on;
        for (int j = tid; j < on; j += THREADS_A) {
            const double a = gr[j], b = gi[j], x = zr[j], y = zi[j], h = hs[j];
            v[0] = fma(h, a, v[0]);
            v[1] = fma(h, b, v[1]);
            v[2] = fma(a, x, fma(-b, y, v[2]));
            v[3] = fma(a, y, fma(b, x, v[3]));
        }
    }
    const double* cr = c + (long long)w * naux;
    const double* ci = c + (long long)(nw + w) * naux;
    for (int g = tid; g < naux; g += THREADS_A) {
        const double a = c_scale * cr[g], b = c_scale * ci[g];
        v[4] = fma(a, a, fma(-b, b, v[4]));
        v[5] = fma(2.0 * a, b, v[5]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][tid] = v[k];
    __syncthreads();
    for (int h = THREADS_A / 2; h > 0; h >>= 1) {
        if (tid < h) {
#pragma unroll
            for (int k = 0; k < 6; ++k) red[k][tid] += red[k][tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const double e1r = red[0][0], e1i = red[1][0], xr = red[2][0], xi = red[3][0];
        const double qr = red[4][0], qi = red[5][0];
        if (nspin == 1)
            E[w] = make_double2(h0 + 2.0 * e1r + 2.0 * qr - xr, 2.0 * e1i + 2.0 * qi - xi);
        else
            E[w] = make_double2(h0 + e1r + 0.5 * (qr - xr), e1i + 0.5 * (qi - xi));
    }
}

__global__ __launch_bounds__(THREADS_A) void energy_assemble_kernel(int nw, int on, int nspin, int naux, double h0,
                                                                    double c_scale, const double* __restrict__ G,
                                                                    const double* __restrict__ Z,
                                                                    const double* __restrict__ c,
                                                                    const double* __restrict__ h1,
                                                                    double2* __restrict__ E) {
    __shared__ double red[6][THREADS_A];
    energy_assemble_walker(blockIdx.x, nw, on, nspin, naux, h0, c_scale, G, Z, c, h1, E, red);
}

// one workgroup per walker of every system (restricted walkers, one sector):
// G = greens, Z = exch_z, c = fb / 2 (the step's force bias, alpha 2), rot_h1,
// h0, energy from the system's descriptor
__global__ __launch_bounds__(THREADS_A) void energy_assemble_grouped_kernel(int count,
                                                                            const afq_system_step* __restrict__ S) {
    __shared__ double red[6][THREADS_A];
    const int b = blockIdx.x;
    const int si = afq::find_group(count, b, [&](int i) { return S[i].walker0; });
    const afq_system_step& d = S[si];
    energy_assemble_walker(b - d.walker0, d.nw, d.o * d.n, 1, d.naux, d.h0, 0.5, d.greens, d.exch_z, d.fb,
                           d.rot_h1, (double2*)d.energy, red);
}

}  // namespace

// Declared in afqmc_hip.h.
extern "C" int afq_energy_assemble(int nw, int on, int nspin, int naux, double h0, double c_scale,
                                   const double* greens, const double* z, const double* c, const double* rot_h1,
                                   double* energy, void* stream) {
    AFQ_REQUIRE(nw >= 0 && on > 0 && naux > 0 && (nspin == 1 || nspin == 2),
                "afq_energy_assemble: bad sizes nw=%d on=%d nspin=%d naux=%d", nw, on, nspin, naux);
    if (nw == 0) return 0;
    AFQ_REQUIRE(greens && z && c && rot_h1 && energy, "afq_energy_assemble: null pointer");
    hipLaunchKernelGGL(energy_assemble_kernel, dim3((unsigned)nw), dim3(THREADS_A), 0, (hipStream_t)stream, nw, on,
                       nspin, naux, h0, c_scale, greens, z, c, rot_h1, (double2*)energy);
    return afq::check_launch("afq_energy_assemble");
}

extern "C" int afq_energy_assemble_grouped(int count, const afq_system_step* sys, int total_walkers,
                                           void* stream) {
    AFQ_REQUIRE(count > 0 && total_walkers > 0, "afq_energy_assemble_grouped: count=%d total_walkers=%d", count,
                total_walkers);
    AFQ_REQUIRE(sys, "afq_energy_assemble_grouped: null pointer");
    hipLaunchKernelGGL(energy_assemble_grouped_kernel, dim3((unsigned)total_walkers), dim3(THREADS_A), 0,
                       (hipStream_t)stream, count, sys);
    return afq::check_launch("afq_energy_assemble_grouped");
}
