// Local energy (mixed estimator) per walker, fp64.
//
// One 256-thread workgroup per walker.  G^T (n x o, split planes) of one spin
// sector sits in LDS as the MFMA B operand; wave wv handles Cholesky vectors
// g = wv, wv+4, ...: f_g = rot_chol_g G^T (o x o) is accumulated in MFMA
// registers from rot_chol_g streamed out of L2 (all walkers walk g in the same
// order), then reduced on the spot:
//   Coulomb  c_g = tr f_g      -- diagonal lanes + wave shuffle sum,
//   exchange tr(f_g f_g)       -- transpose through a per-wave LDS tile,
// so the (naux, o, o) intermediate of meas/rhf.py:45-47 is never formed.
#include "afqmc_hip.h"
#include "common.h"

namespace {

constexpr int THREADS = 256;

template <int NT>
__global__ __launch_bounds__(THREADS, 1) void energy_kernel(
    int nw, int n, int o, int nspin, int naux, double h0, const double* __restrict__ greens,
    const double* __restrict__ rotp, int ldr, const double* __restrict__ rot_h1,
    double2* __restrict__ eout) {
    extern __shared__ double lds[];
    constexpr int TC = NT * 16;
    constexpr int XP = TC + 1;
    constexpr int SP = TC + 1;
    __shared__ double red[4][6];
    const int NP = afq::round_up(n, 16);
    double* Xr = lds;
    double* Xi = Xr + NP * XP;
    double* S = Xi + NP * XP;
    double2* cs = reinterpret_cast<double2*>(S + 4 * TC * SP);

    const int w = blockIdx.x;
    const long long ldg = (long long)nspin * o * n;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    double* Sw = S + wave * TC * SP;

    double e1r = 0.0, e1i = 0.0;  // one-body partial sums
    double exr = 0.0, exi = 0.0;  // exchange partial sums (per lane)
    double c2r = 0.0, c2i = 0.0;  // Coulomb sums (uniform within a wave)

    for (int s = 0; s < nspin; ++s) {
        const double* gre = greens + (long long)w * ldg + (long long)s * o * n;
        const double* gim = gre + (long long)nw * ldg;
        __syncthreads();
        for (int e = tid; e < NP * TC; e += THREADS) {
            const int j = e / NP;
            const int p = e - j * NP;
            double vr = 0.0, vi = 0.0;
            if (p < n && j < o) {
                vr = gre[j * n + p];
                vi = gim[j * n + p];
            }
            Xr[p * XP + j] = vr;
            Xi[p * XP + j] = vi;
        }
        const double* rh = rot_h1 + (long long)s * o * n;
        for (int e = tid; e < o * n; e += THREADS) {
            const double h = rh[e];
            e1r += gre[e] * h;
            e1i += gim[e] * h;
        }
        __syncthreads();

        const double* R = rotp + (long long)s * naux * TC * ldr;
        const int nkb = afq::ceil_div(n, 16);
        for (int g0 = 0; g0 < naux; g0 += 4) {
            const int g = g0 + wave;
            const bool active = g < naux;
            f64x4 F[NT][NT][2];
#pragma unroll
            for (int a = 0; a < NT; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b) {
                    F[a][b][0] = f64x4{0.0, 0.0, 0.0, 0.0};
                    F[a][b][1] = f64x4{0.0, 0.0, 0.0, 0.0};
                }
            if (active) {
                const double* A = R + (long long)g * TC * ldr;
                for (int kb = 0; kb < nkb; ++kb) {
                    const int k0 = kb * 16 + lk;
                    double br[4][NT], bi[4][NT];
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                        for (int tj = 0; tj < NT; ++tj) {
                            br[kk][tj] = Xr[(k0 + kk) * XP + tj * 16 + lr];
                            bi[kk][tj] = Xi[(k0 + kk) * XP + tj * 16 + lr];
                        }
#pragma unroll
                    for (int ti = 0; ti < NT; ++ti) {
                        double a[4] = {0.0, 0.0, 0.0, 0.0};
                        if (k0 < n) {
                            const double4 v = *reinterpret_cast<const double4*>(
                                A + (long long)(ti * 16 + lr) * ldr + k0);
                            a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
                        }
#pragma unroll
                        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                            for (int tj = 0; tj < NT; ++tj) {
                                F[ti][tj][0] = afq::mfma_f64(a[kk], br[kk][tj], F[ti][tj][0]);
                                F[ti][tj][1] = afq::mfma_f64(a[kk], bi[kk][tj], F[ti][tj][1]);
                            }
                    }
                }
            }
            // Coulomb: c_g = tr f_g  (element r of a diagonal tile sits on the
            // diagonal when its row, 4*r + lane/16, equals its column lane%16)
            double cr = 0.0, ci = 0.0;
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (lr == afq::mfma16<double>::row(lane, r)) {
                        cr += F[t][t][0][r];
                        ci += F[t][t][1][r];
                    }
            cr = afq::wave_sum(cr);
            ci = afq::wave_sum(ci);
            if (active) {
                if (nspin == 1) {
                    c2r += cr * cr - ci * ci;
                    c2i += 2.0 * cr * ci;
                } else if (s == 0) {
                    if (lane == 0) cs[g] = make_double2(cr, ci);
                } else {
                    const double2 cu = cs[g];
                    const double tr = cu.x + cr;
                    const double ti = cu.y + ci;
                    c2r += tr * tr - ti * ti;
                    c2i += 2.0 * tr * ti;
                }
            }
            // exchange: sum_ij f_ij f_ji, real plane then imaginary plane
#pragma unroll
            for (int pl = 0; pl < 2; ++pl) {
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = ti * 16 + afq::mfma16<double>::row(lane, r);
                            const int col = tj * 16 + lr;
                            Sw[row * SP + col] = F[ti][tj][pl][r];
                        }
                __syncthreads();
#pragma unroll
                for (int ti = 0; ti < NT; ++ti)
#pragma unroll
                    for (int tj = 0; tj < NT; ++tj)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int row = ti * 16 + afq::mfma16<double>::row(lane, r);
                            const int col = tj * 16 + lr;
                            const double t = Sw[col * SP + row];
                            if (pl == 0) {
                                exr += F[ti][tj][0][r] * t;
                                exi += F[ti][tj][1][r] * t;
                            } else {
                                exr -= F[ti][tj][1][r] * t;
                                exi += F[ti][tj][0][r] * t;
                            }
                        }
                __syncthreads();
            }
        }
    }

    e1r = afq::wave_sum(e1r);
    e1i = afq::wave_sum(e1i);
    exr = afq::wave_sum(exr);
    exi = afq::wave_sum(exi);
    if (lane == 0) {
        red[wave][0] = e1r;
        red[wave][1] = e1i;
        red[wave][2] = exr;
        red[wave][3] = exi;
        red[wave][4] = c2r;
        red[wave][5] = c2i;
    }
    __syncthreads();
    if (tid == 0) {
        double v[6] = {0, 0, 0, 0, 0, 0};
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 6; ++k) v[k] += red[q][k];
        double er, ei;
        if (nspin == 1) {
            er = h0 + 2.0 * v[0] + 2.0 * v[4] - v[2];
            ei = 2.0 * v[1] + 2.0 * v[5] - v[3];
        } else {
            er = h0 + v[0] + 0.5 * (v[4] - v[2]);
            ei = v[1] + 0.5 * (v[5] - v[3]);
        }
        eout[w] = make_double2(er, ei);
    }
}

template <int NT>
int launch_energy(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
                  const double* rotp, int ldr, const double* rot_h1, double* energy,
                  hipStream_t st) {
    constexpr int TC = NT * 16;
    const int NP = afq::round_up(n, 16);
    const size_t lds = sizeof(double) * ((size_t)2 * NP * (TC + 1) + (size_t)4 * TC * (TC + 1) +
                                         (nspin == 2 ? (size_t)2 * naux : 0));
    AFQ_REQUIRE(lds + 4 * 6 * sizeof(double) <= 160 * 1024,
                "afq_energy: n=%d o=%d naux=%d needs more than 160 KB of LDS", n, o, naux);
    if (lds > 64 * 1024) {
        (void)hipFuncSetAttribute((const void*)energy_kernel<NT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipGetLastError();
    }
    hipLaunchKernelGGL(energy_kernel<NT>, dim3((unsigned)nw), dim3(THREADS), lds, st, nw, n, o, nspin,
                       naux, h0, greens, rotp, ldr, rot_h1, (double2*)energy);
    return afq::check_launch("afq_energy");
}

}  // namespace

extern "C" int afq_energy(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
                          const double* rot_chol_pad, int ldr, const double* rot_h1, double* energy,
                          void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n && naux > 0, "afq_energy: bad sizes");
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "afq_energy: nspin must be 1 or 2");
    AFQ_REQUIRE(ldr >= n && ldr % 4 == 0, "afq_energy: ldr=%d must be >= n and a multiple of 4", ldr);
    AFQ_REQUIRE(greens && rot_chol_pad && rot_h1 && energy, "afq_energy: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (o <= 16) return launch_energy<1>(nw, n, o, nspin, naux, h0, greens, rot_chol_pad, ldr, rot_h1, energy, st);
    if (o <= 32) return launch_energy<2>(nw, n, o, nspin, naux, h0, greens, rot_chol_pad, ldr, rot_h1, energy, st);
    afq::set_error("afq_energy: o=%d exceeds the supported maximum of 32", o);
    return -1;
}
