// Fused Trotter step  w' = E . sum_{k<=K} (i sqrt_dt V)^k / k! . E w, in fp64
// (plain step) or fp32 (mixed step: E, V and the Taylor terms in fp32, walkers
// read and written as complex128).
//
// One 256-thread workgroup per (walker, spin).  The running Taylor term T
// (n x o complex, split planes) lives in LDS (2 * ceil16(n) * 16*NC doubles;
// 106 KB at n=200, o=32 in fp64); the running sum lives in MFMA accumulators.  Every
// product is  P = A . T  with A either the real one-body propagator E or the
// walker's complex VHS, streamed from global memory (L2/MALL) straight into
// MFMA A-operand registers, one (k-block, row tile) step ahead of use.  Wave wv owns the row
// tiles wv, wv+4, ...; each lane feeds four consecutive k of one row to four
// v_mfma_{f64,f32}_16x16x4 (the B operand from LDS uses the same k order).
//
// Sequence per workgroup (prop/chol_afqmc_ops.py:173-183, prop/utils.py:20-31):
//   T = E w;  out = T;  for k: T = (i sqrt_dt / k) V T, out += T;  w' = E out.
#include "afqmc_hip.h"
#include "common.h"

namespace {

// P[i][plane] = A . T for the row tiles (rg + 4 i) and column tile ct of this
// wave.  The A operand of (k-block, row tile) step q+1 is loaded while the
// MFMAs of step q run.
template <typename R, int NC, int RPW, bool CPLX>
__device__ __forceinline__ void tile_product(const R* __restrict__ Ar, const R* __restrict__ Ai,
                                             int lda, int n, int ntile, const R* Tr, const R* Ti,
                                             int rg, int ct, int lane,
                                             typename afq::mfma16<R>::acc (&P)[RPW][2]) {
    using MF = afq::mfma16<R>;
    using V4 = typename afq::vec4<R>::type;
    constexpr int TC = NC * 16;
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        P[i][0] = MF::zero();
        P[i][1] = MF::zero();
    }
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    const int nkb = afq::ceil_div(n, 16);
    const int col = ct * 16 + lr;

    // Branch-free operand load: out-of-range rows / k-blocks read a clamped,
    // valid address and are zeroed by a select.
    auto load_a = [&](int kb, int i, R (&xr)[4], R (&xi)[4]) {
        const int k0 = kb * 16 + lk;
        const int rt = rg + 4 * i;
        const int row = rt * 16 + lr;
        const bool ok = rt < ntile && row < n && k0 < n;
        const int rowc = ok ? row : 0;
        const int k0c = ok ? k0 : 0;
        const V4 v = *reinterpret_cast<const V4*>(Ar + (long long)rowc * lda + k0c);
        xr[0] = ok ? v.x : R(0); xr[1] = ok ? v.y : R(0); xr[2] = ok ? v.z : R(0); xr[3] = ok ? v.w : R(0);
        if (CPLX) {
            const V4 u = *reinterpret_cast<const V4*>(Ai + (long long)rowc * lda + k0c);
            xi[0] = ok ? u.x : R(0); xi[1] = ok ? u.y : R(0); xi[2] = ok ? u.z : R(0); xi[3] = ok ? u.w : R(0);
        }
    };

    R ar[4], ai[4];
    load_a(0, 0, ar, ai);
#pragma unroll 1
    for (int kb = 0; kb < nkb; ++kb) {
        const int k0 = kb * 16 + lk;
        R br[4], bi[4], nbi[4];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            br[kk] = Tr[(k0 + kk) * TC + col];
            bi[kk] = Ti[(k0 + kk) * TC + col];
            nbi[kk] = -bi[kk];
        }
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            R nr[4], ni[4];
            const int nkb_i = (i + 1 < RPW) ? kb : kb + 1;
            const int ni_i = (i + 1 < RPW) ? i + 1 : 0;
            // the last step prefetches a clamped (discarded) address
            load_a(nkb_i < nkb ? nkb_i : kb, ni_i, nr, ni);
            // tiles past the last row tile multiply zeros (no divergent accumulators)
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                P[i][0] = MF::fma(ar[kk], br[kk], P[i][0]);
                P[i][1] = MF::fma(ar[kk], bi[kk], P[i][1]);
                if (CPLX) {
                    // re -= Vi Ti ;  im += Vi Tr
                    P[i][0] = MF::fma(ai[kk], nbi[kk], P[i][0]);
                    P[i][1] = MF::fma(ai[kk], br[kk], P[i][1]);
                }
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                ar[kk] = nr[kk];
                if (CPLX) ai[kk] = ni[kk];
            }
        }
    }
}

template <typename R, int NC, int RPW>
__device__ __forceinline__ void store_T(R* Tr, R* Ti, int rg, int ct, int lane, int ntile,
                                        const typename afq::mfma16<R>::acc (&X)[RPW][2]) {
    constexpr int TC = NC * 16;
    const int col = ct * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int rt = rg + 4 * i;
        if (rt < ntile) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rt * 16 + afq::mfma16<R>::row(lane, r);
                Tr[row * TC + col] = X[i][0][r];
                Ti[row * TC + col] = X[i][1][r];
            }
        }
    }
}

// 4 * NC waves: wave -> (row group rg = wave % 4, column tile ct = wave / 4).
template <typename R, int NC, int RPW>
__global__ __launch_bounds__(256 * NC, 1) void trotter_kernel(
    int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
    const R* __restrict__ E, const R* __restrict__ vhs, long long vplane,
    const double2* __restrict__ win_up, const double2* __restrict__ win_dn,
    double2* __restrict__ wout_up, double2* __restrict__ wout_dn, int pair_map, int ngrp) {
    using ACC = typename afq::mfma16<R>::acc;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    R* lds = reinterpret_cast<R*>(lds_raw);
    constexpr int TC = NC * 16;
    constexpr int THREADS = 256 * NC;
    const int NP = afq::round_up(n, 16);
    const int ntile = NP / 16;
    R* Tr = lds;
    R* Ti = lds + NP * TC;

    // o > TC: the workgroups of one (walker, spin) take consecutive groups of
    // TC columns (the Taylor recursion never mixes columns) and run back to
    // back, reading the walker's VHS from L2 after the first.
    const int grp = blockIdx.x % ngrp;
    const int c0 = grp * TC;
    // Both spins of a walker go to the same XCD, close together in time, so
    // the second reads the walker's VHS from L2.
    int w, s;
    const int b = blockIdx.x / ngrp;
    if (nspin == 1) {
        w = b;
        s = 0;
    } else if (pair_map) {
        w = (b / 16) * 8 + (b % 8);
        s = (b / 8) % 2;
    } else {
        w = b / 2;
        s = b % 2;
    }
    const long long woff = (long long)w * n * o;
    const double2* Win = (s == 0 ? win_up : win_dn) + woff;
    double2* Wout = (s == 0 ? wout_up : wout_dn) + woff;
    const R* Vr = vhs + (long long)w * n * ldv;
    const R* Vi = Vr + vplane;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int rg = wave & 3;
    const int ct = wave >> 2;

    for (int e = tid; e < NP * TC; e += THREADS) {
        const int p = e / TC;
        const int j = c0 + (e - p * TC);
        const double2 x = (p < n && j < o) ? Win[p * o + j] : make_double2(0.0, 0.0);
        Tr[e] = (R)x.x;
        Ti[e] = (R)x.y;
    }
    __syncthreads();

    ACC P[RPW][2];
    ACC out[RPW][2];

    // T = E w
    tile_product<R, NC, RPW, false>(E, nullptr, ldv, n, ntile, Tr, Ti, rg, ct, lane, P);
    __syncthreads();
    store_T<R, NC, RPW>(Tr, Ti, rg, ct, lane, ntile, P);
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        out[i][0] = P[i][0];
        out[i][1] = P[i][1];
    }
    __syncthreads();

#pragma unroll 1
    for (int k = 1; k <= n_terms; ++k) {
        tile_product<R, NC, RPW, true>(Vr, Vi, ldv, n, ntile, Tr, Ti, rg, ct, lane, P);
        const R ck = (R)(sqrt_dt / (double)k);
        // (i ck) * P
#pragma unroll
        for (int i = 0; i < RPW; ++i) {
            const ACC re = -ck * P[i][1];
            const ACC im = ck * P[i][0];
            P[i][0] = re;
            P[i][1] = im;
            out[i][0] += re;
            out[i][1] += im;
        }
        __syncthreads();
        store_T<R, NC, RPW>(Tr, Ti, rg, ct, lane, ntile, P);
        __syncthreads();
    }

    // w' = E out
    store_T<R, NC, RPW>(Tr, Ti, rg, ct, lane, ntile, out);
    __syncthreads();
    tile_product<R, NC, RPW, false>(E, nullptr, ldv, n, ntile, Tr, Ti, rg, ct, lane, P);
    const int col = c0 + ct * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
        const int rt = rg + 4 * i;
        if (rt < ntile && col < o) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = rt * 16 + afq::mfma16<R>::row(lane, r);
                if (row < n) Wout[row * o + col] = make_double2((double)P[i][0][r], (double)P[i][1][r]);
            }
        }
    }
}

template <typename R, int NC, int RPW>
int launch_trotter(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                   const R* E, const R* vhs, const double2* wu, const double2* wd,
                   double2* ou, double2* od, hipStream_t st) {
    const int NP = afq::round_up(n, 16);
    const size_t lds = (size_t)2 * NP * NC * 16 * sizeof(R);
    const int pair_map = (nspin == 2 && nw % 8 == 0) ? 1 : 0;
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_kernel<R, NC, RPW>);
    const int ngrp = afq::ceil_div(o, NC * 16);
    hipLaunchKernelGGL((trotter_kernel<R, NC, RPW>), dim3((unsigned)(nw * nspin * ngrp)), dim3(256 * NC), lds,
                       st, nw, n, o, nspin, ldv, n_terms, sqrt_dt, E, vhs,
                       (long long)nw * n * ldv, wu, wd, ou, od, pair_map, ngrp);
    return afq::check_launch("afq_trotter");
}

// ---------------------------------------------------------------------------
// fp32 Trotter step with the walker's VHS resident on chip for all Taylor
// terms.  One wave per 16-row tile of V (ceil16(n)/16 waves, 13 at n = 200):
// each wave loads its V rows once; k-blocks (16 k) 0..KRB-1 stay in registers
// as MFMA A fragments, the remaining k-blocks sit in LDS in fragment order
// (lane-contiguous, conflict-free).  The Taylor recursion never mixes columns
// of T, so each workgroup owns one 16-column tile of one (walker, spin): T
// (n x 16, split planes) is the B operand in LDS and every wave computes its
// whole row tile, so no partial sums cross waves.  The workgroups of one
// walker run back to back on one XCD and share V through L2; V crosses HBM
// once per walker instead of once per Taylor term.
// ---------------------------------------------------------------------------
// One 16-column tile (first column c0) of one walker: Win -> Wout with the
// walker's V planes Vr / Vi.  LDS holds T for NPL >= round16(n) rows; waves
// beyond round16(n) / 16 (a workgroup sized for a larger system) compute on
// zero rows, write only LDS rows no other wave reads, and store nothing.
template <int KRB>
__device__ __forceinline__ void trotter_resident_body(
    int n, int o, int ldv, int n_terms, double sqrt_dt, const float* __restrict__ E,
    const float* __restrict__ Vr, const float* __restrict__ Vi, const double2* __restrict__ Win,
    double2* __restrict__ Wout, int c0, int NPL, unsigned char* lds_raw) {
    constexpr int NC = 1;
    using MF = afq::mfma16<float>;
    using ACC = MF::acc;
    constexpr int TC = NC * 16;
    const int NP = afq::round_up(n, 16);
    const int nkb = NP / 16;
    const int nkl = nkb > KRB ? nkb - KRB : 0;
    float* Tr = reinterpret_cast<float*>(lds_raw);
    float* Ti = Tr + NPL * TC;
    float* VL = Ti + NPL * TC;  // [wave][nkl][4 kk][2 planes][64 lanes]

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    const int rt = wave;
    const int row = rt * 16 + lr;

    // branch-free fragment load of 4 consecutive k of this lane's row
    auto load4 = [&](const float* A, int kb, float (&x)[4]) {
        const int k0 = kb * 16 + lk;
        const bool ok = row < n && k0 < n;
        const float4 v = *reinterpret_cast<const float4*>(A + (long long)(ok ? row : 0) * ldv + (ok ? k0 : 0));
        x[0] = ok ? v.x : 0.f; x[1] = ok ? v.y : 0.f; x[2] = ok ? v.z : 0.f; x[3] = ok ? v.w : 0.f;
    };

    for (int e = tid; e < NP * TC; e += nthreads) {
        const int p = e / TC;
        const int j = c0 + (e - p * TC);
        const double2 x = (p < n && j < o) ? Win[p * o + j] : make_double2(0.0, 0.0);
        Tr[e] = (float)x.x;
        Ti[e] = (float)x.y;
    }
    __syncthreads();

    ACC P[NC][2], out[NC][2];

    // P = E . T (E real, from L2): all of this wave's E fragments are loaded
    // first (nkb <= 16), then consumed, so the loads overlap
    auto e_product = [&]() {
        float a[16][4];
#pragma unroll
        for (int kb = 0; kb < 16; ++kb)
            if (kb < nkb) load4(E, kb, a[kb]);
#pragma unroll
        for (int c = 0; c < NC; ++c) P[c][0] = P[c][1] = MF::zero();
#pragma unroll
        for (int kb = 0; kb < 16; ++kb) {
            if (kb < nkb) {
                const int k0 = kb * 16 + lk;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        const int col = c * 16 + lr;
                        P[c][0] = MF::fma(a[kb][kk], Tr[(k0 + kk) * TC + col], P[c][0]);
                        P[c][1] = MF::fma(a[kb][kk], Ti[(k0 + kk) * TC + col], P[c][1]);
                    }
            }
        }
    };
    auto store_T = [&](const ACC (&X)[NC][2]) {
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = rt * 16 + MF::row(lane, r);
                Tr[rr * TC + c * 16 + lr] = X[c][0][r];
                Ti[rr * TC + c * 16 + lr] = X[c][1][r];
            }
    };
    // complex V . T step for one (k-block, kk) with A = (ar, ai)
    auto vt_step = [&](int k, float ar, float ai) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int col = c * 16 + lr;
            const float br = Tr[k * TC + col];
            const float bi = Ti[k * TC + col];
            P[c][0] = MF::fma(ar, br, P[c][0]);
            P[c][1] = MF::fma(ar, bi, P[c][1]);
            P[c][0] = MF::fma(ai, -bi, P[c][0]);
            P[c][1] = MF::fma(ai, br, P[c][1]);
        }
    };

    // T = E w, before V occupies the registers
    e_product();

    float vr[KRB][4], vi[KRB][4];
#pragma unroll
    for (int kb = 0; kb < KRB; ++kb) {
        load4(Vr, kb < nkb ? kb : 0, vr[kb]);
        load4(Vi, kb < nkb ? kb : 0, vi[kb]);
    }
    for (int kb = KRB; kb < nkb; ++kb) {
        float xr[4], xi[4];
        load4(Vr, kb, xr);
        load4(Vi, kb, xi);
        float* dst = VL + ((size_t)(wave * nkl + (kb - KRB)) * 8) * 64 + lane;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            dst[(kk * 2 + 0) * 64] = xr[kk];
            dst[(kk * 2 + 1) * 64] = xi[kk];
        }
    }

    __syncthreads();
    store_T(P);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        out[c][0] = P[c][0];
        out[c][1] = P[c][1];
    }
    __syncthreads();

#pragma unroll 1
    for (int k = 1; k <= n_terms; ++k) {
#pragma unroll
        for (int c = 0; c < NC; ++c) P[c][0] = P[c][1] = MF::zero();
#pragma unroll
        for (int kb = 0; kb < KRB; ++kb) {
            if (kb < nkb) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) vt_step(kb * 16 + lk + kk, vr[kb][kk], vi[kb][kk]);
            }
        }
        for (int kb = KRB; kb < nkb; ++kb) {
            const float* src = VL + ((size_t)(wave * nkl + (kb - KRB)) * 8) * 64 + lane;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                vt_step(kb * 16 + lk + kk, src[(kk * 2 + 0) * 64], src[(kk * 2 + 1) * 64]);
        }
        const float ck = (float)(sqrt_dt / (double)k);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const ACC re = -ck * P[c][1];
            const ACC im = ck * P[c][0];
            P[c][0] = re;
            P[c][1] = im;
            out[c][0] += re;
            out[c][1] += im;
        }
        __syncthreads();
        store_T(P);
        __syncthreads();
    }

    store_T(out);
    __syncthreads();
    e_product();
    {
        const int col = c0 + lr;
        if (col < o) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = rt * 16 + MF::row(lane, r);
                if (rr < n) Wout[rr * o + col] = make_double2((double)P[0][0][r], (double)P[0][1][r]);
            }
        }
    }
}

template <int KRB>
__global__ __launch_bounds__(1024, 1) void trotter_resident_kernel(
    int nw, int n, int o, int nspin, int nct, int ldv, int n_terms, double sqrt_dt,
    const float* __restrict__ E, const float* __restrict__ vhs, long long vplane,
    const double2* __restrict__ win_up, const double2* __restrict__ win_dn,
    double2* __restrict__ wout_up, double2* __restrict__ wout_dn, int group_map) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    // block -> (walker, spin, column tile); the G = nspin * nct blocks of a
    // walker land on one XCD (blocks b, b+8, ... share an XCD) back to back
    const int G = nspin * nct;
    const int b = blockIdx.x;
    int w, sub;
    if (group_map) {
        w = (b / (8 * G)) * 8 + (b % 8);
        sub = (b / 8) % G;
    } else {
        w = b / G;
        sub = b % G;
    }
    const int s = sub / nct;
    const int c0 = (sub - s * nct) * 16;  // first walker column of this tile
    const long long woff = (long long)w * n * o;
    const float* Vr = vhs + (long long)w * n * ldv;
    trotter_resident_body<KRB>(n, o, ldv, n_terms, sqrt_dt, E, Vr, Vr + vplane,
                               (s == 0 ? win_up : win_dn) + woff, (s == 0 ? wout_up : wout_dn) + woff,
                               c0, afq::round_up(n, 16), lds_raw);
}

// Every system's walkers in one launch: workgroup -> (system, walker, column
// tile) by the systems' tblock0 prefix sums, after an XCD-contiguous remap so
// that the column tiles of one walker share an XCD (and its V in L2).  The
// workgroup has 64 * round16(max_n) / 16 threads and LDS for max_n rows.
template <int KRB>
__global__ __launch_bounds__(1024, 1) void trotter_resident_grouped_kernel(
    int count, int total, int NPL, int n_terms, const afq_system_step* __restrict__ S) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int b = afq::xcd_contiguous(blockIdx.x, total);
    const int si = afq::find_group(count, b, [&](int i) { return S[i].tblock0; });
    const afq_system_step& d = S[si];
    const int n = d.n, o = d.o, ldv = d.ldv;
    const int nct = afq::ceil_div(o, 16);
    const int lb = b - d.tblock0;
    const int w = lb / nct;
    const int c0 = (lb - w * nct) * 16;
    const long long woff = (long long)w * n * o;
    const float* Vr = d.vhs + (long long)w * n * ldv;
    trotter_resident_body<KRB>(n, o, ldv, n_terms, d.sqrt_dt, d.exp_h1_half, Vr,
                               Vr + (long long)d.nw * n * ldv, (const double2*)d.walkers_in + woff,
                               (double2*)d.walkers_out + woff, c0, NPL, lds_raw);
}

constexpr int RESIDENT_KRB = 9;

size_t resident_lds(int n) {
    const int NP = afq::round_up(n, 16);
    const int nkb = NP / 16;
    const int nkl = nkb > RESIDENT_KRB ? nkb - RESIDENT_KRB : 0;
    return sizeof(float) * ((size_t)2 * NP * 16 + (size_t)(NP / 16) * nkl * 8 * 64);
}

int launch_trotter_resident(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                            const float* E, const float* vhs, const double2* wu, const double2* wd,
                            double2* ou, double2* od, hipStream_t st) {
    const int NP = afq::round_up(n, 16);
    const int nct = afq::ceil_div(o, 16);
    const size_t lds = resident_lds(n);
    const int group_map = (nw % 8 == 0) ? 1 : 0;
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_resident_kernel<RESIDENT_KRB>);
    hipLaunchKernelGGL((trotter_resident_kernel<RESIDENT_KRB>), dim3((unsigned)(nw * nspin * nct)),
                       dim3(64 * (NP / 16)), lds, st, nw, n, o, nspin, nct, ldv, n_terms, sqrt_dt, E,
                       vhs, (long long)nw * n * ldv, wu, wd, ou, od, group_map);
    return afq::check_launch("afq_trotter_f32");
}

// ---------------------------------------------------------------------------
// fp64 Trotter step with the walker's VHS resident on chip, for n <= 128.
// One workgroup per (walker, column chunk) where the chunk runs over the
// walker's columns of both spins laid side by side (spin-up columns 0..o-1,
// then spin-down): the spins share V, so V crosses HBM once per walker and
// the MFMA column tiles are filled across the spin boundary (o = 20 per spin:
// 40 columns in three 16-column tiles instead of two tiles per spin).  One
// wave per 16-row tile of V; k-blocks 0..KRB-1 of its rows stay in
// registers as MFMA A fragments, the rest in LDS in fragment order.  T
// (ceil16(n) x 16*NCT, split planes) is the B operand in LDS; each wave
// computes its row tile for every column tile, so no partial sums cross waves.
// ---------------------------------------------------------------------------
constexpr int F64_RES_MAX_NP = 128;
constexpr int F64_RES_MAX_COLS = 48;

// register-resident k-blocks of V: three column tiles leave room for six
__host__ __device__ constexpr int f64_res_krb(int nct) { return nct >= 3 ? 6 : 7; }

size_t f64_resident_lds(int n, int nct) {
    const int NP = afq::round_up(n, 16);
    const int nkb = NP / 16;
    const int krb = f64_res_krb(nct);
    const int nkl = nkb > krb ? nkb - krb : 0;
    return sizeof(double) * ((size_t)2 * NP * 16 * nct + (size_t)(NP / 16) * nkl * 8 * 64);
}

template <int NCT>
__global__ __launch_bounds__(512, 1) void trotter_f64_resident_kernel(
    int nw, int n, int o, int nspin, int chunk, int ngrp, int ldv, int n_terms, double sqrt_dt,
    const double* __restrict__ E, const double* __restrict__ vhs, long long vplane,
    const double2* __restrict__ win_up, const double2* __restrict__ win_dn,
    double2* __restrict__ wout_up, double2* __restrict__ wout_dn) {
    constexpr int KRB = f64_res_krb(NCT);
    constexpr int TC = 16 * NCT;
    using MF = afq::mfma16<double>;
    using ACC = MF::acc;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int NP = afq::round_up(n, 16);
    const int nkb = NP / 16;
    const int nkl = nkb > KRB ? nkb - KRB : 0;
    double* Tr = reinterpret_cast<double*>(lds_raw);
    double* Ti = Tr + NP * TC;
    double* VL = Ti + NP * TC;  // [wave][nkl][4 kk][2 planes][64 lanes]

    // the ngrp chunks of one walker are consecutive blocks
    const int w = blockIdx.x / ngrp;
    const int j0 = (blockIdx.x - w * ngrp) * chunk;  // first merged column
    const int ncols = nspin * o;
    const long long woff = (long long)w * n * o;
    const double* Vr = vhs + (long long)w * n * ldv;
    const double* Vi = Vr + vplane;

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int lane = tid & 63;
    const int rt = tid >> 6;  // row tile of this wave
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    const int row = rt * 16 + lr;

    auto load4 = [&](const double* A, int kb, double (&x)[4]) {
        const int k0 = kb * 16 + lk;
        const bool ok = row < n && k0 < n;
        const double4 v = *reinterpret_cast<const double4*>(A + (long long)(ok ? row : 0) * ldv + (ok ? k0 : 0));
        x[0] = ok ? v.x : 0.0; x[1] = ok ? v.y : 0.0; x[2] = ok ? v.z : 0.0; x[3] = ok ? v.w : 0.0;
    };

    // merged column t of this chunk -> (spin, walker column); -1 past the end
    auto col_of = [&](int t, int& s, int& c) {
        const int j = j0 + t;
        if (t >= chunk || j >= ncols) return false;
        s = j >= o ? 1 : 0;
        c = j - s * o;
        return true;
    };

    for (int e = tid; e < NP * TC; e += nthreads) {
        const int p = e / TC;
        const int t = e - p * TC;
        int s, c;
        double2 x = make_double2(0.0, 0.0);
        if (p < n && col_of(t, s, c)) x = (s == 0 ? win_up : win_dn)[woff + (long long)p * o + c];
        Tr[e] = x.x;
        Ti[e] = x.y;
    }
    __syncthreads();

    ACC P[NCT][2], out[NCT][2];

    // P = E . T, E real from L2; this wave's E fragments are loaded first
    auto e_product = [&]() {
        double a[F64_RES_MAX_NP / 16][4];
#pragma unroll
        for (int kb = 0; kb < F64_RES_MAX_NP / 16; ++kb)
            if (kb < nkb) load4(E, kb, a[kb]);
#pragma unroll
        for (int c = 0; c < NCT; ++c) P[c][0] = P[c][1] = MF::zero();
#pragma unroll
        for (int kb = 0; kb < F64_RES_MAX_NP / 16; ++kb) {
            if (kb < nkb) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int k = kb * 16 + lk + kk;
#pragma unroll
                    for (int c = 0; c < NCT; ++c) {
                        P[c][0] = MF::fma(a[kb][kk], Tr[k * TC + c * 16 + lr], P[c][0]);
                        P[c][1] = MF::fma(a[kb][kk], Ti[k * TC + c * 16 + lr], P[c][1]);
                    }
                }
            }
        }
    };
    auto store_T = [&](const ACC (&X)[NCT][2]) {
#pragma unroll
        for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = rt * 16 + MF::row(lane, r);
                Tr[rr * TC + c * 16 + lr] = X[c][0][r];
                Ti[rr * TC + c * 16 + lr] = X[c][1][r];
            }
    };
    auto vt_step = [&](int k, double ar, double ai) {
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const double br = Tr[k * TC + c * 16 + lr];
            const double bi = Ti[k * TC + c * 16 + lr];
            P[c][0] = MF::fma(ar, br, P[c][0]);
            P[c][1] = MF::fma(ar, bi, P[c][1]);
            P[c][0] = MF::fma(ai, -bi, P[c][0]);
            P[c][1] = MF::fma(ai, br, P[c][1]);
        }
    };

    // T = E w, before V occupies the registers
    e_product();

    double vr[KRB][4], vi[KRB][4];
#pragma unroll
    for (int kb = 0; kb < KRB; ++kb) {
        load4(Vr, kb < nkb ? kb : 0, vr[kb]);
        load4(Vi, kb < nkb ? kb : 0, vi[kb]);
    }
    for (int kb = KRB; kb < nkb; ++kb) {
        double xr[4], xi[4];
        load4(Vr, kb, xr);
        load4(Vi, kb, xi);
        double* dst = VL + ((size_t)(rt * nkl + (kb - KRB)) * 8) * 64 + lane;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            dst[(kk * 2 + 0) * 64] = xr[kk];
            dst[(kk * 2 + 1) * 64] = xi[kk];
        }
    }

    __syncthreads();
    store_T(P);
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
        out[c][0] = P[c][0];
        out[c][1] = P[c][1];
    }
    __syncthreads();

#pragma unroll 1
    for (int k = 1; k <= n_terms; ++k) {
#pragma unroll
        for (int c = 0; c < NCT; ++c) P[c][0] = P[c][1] = MF::zero();
#pragma unroll
        for (int kb = 0; kb < KRB; ++kb) {
            if (kb < nkb) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) vt_step(kb * 16 + lk + kk, vr[kb][kk], vi[kb][kk]);
            }
        }
        for (int kb = KRB; kb < nkb; ++kb) {
            const double* src = VL + ((size_t)(rt * nkl + (kb - KRB)) * 8) * 64 + lane;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                vt_step(kb * 16 + lk + kk, src[(kk * 2 + 0) * 64], src[(kk * 2 + 1) * 64]);
        }
        const double ck = sqrt_dt / (double)k;
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const ACC re = -ck * P[c][1];
            const ACC im = ck * P[c][0];
            P[c][0] = re;
            P[c][1] = im;
            out[c][0] += re;
            out[c][1] += im;
        }
        __syncthreads();
        store_T(P);
        __syncthreads();
    }

    store_T(out);
    __syncthreads();
    e_product();
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
        int s, cc;
        if (col_of(c * 16 + lr, s, cc)) {
            double2* Wout = (s == 0 ? wout_up : wout_dn) + woff;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = rt * 16 + MF::row(lane, r);
                if (rr < n) Wout[(long long)rr * o + cc] = make_double2(P[c][0][r], P[c][1][r]);
            }
        }
    }
}

// merged-spin column chunking: chunks of at most F64_RES_MAX_COLS columns
void f64_resident_shape(int o, int nspin, int& ngrp, int& chunk, int& nct) {
    const int cols = nspin * o;
    ngrp = afq::ceil_div(cols, F64_RES_MAX_COLS);
    chunk = afq::ceil_div(cols, ngrp);
    nct = afq::ceil_div(chunk, 16);
}

int launch_trotter_f64_resident(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                                const double* E, const double* vhs, const double2* wu, const double2* wd,
                                double2* ou, double2* od, hipStream_t st) {
    const int NP = afq::round_up(n, 16);
    int ngrp, chunk, nct;
    f64_resident_shape(o, nspin, ngrp, chunk, nct);
    const size_t lds = f64_resident_lds(n, nct);
    const dim3 grid((unsigned)(nw * ngrp)), block(64 * (NP / 16));
    const long long vplane = (long long)nw * n * ldv;
#define AFQ_F64_RES(NCT_)                                                                              \
    if (nct == NCT_) {                                                                                 \
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_f64_resident_kernel<NCT_>);     \
        hipLaunchKernelGGL((trotter_f64_resident_kernel<NCT_>), grid, block, lds, st, nw, n, o, nspin, \
                           chunk, ngrp, ldv, n_terms, sqrt_dt, E, vhs, vplane, wu, wd, ou, od);        \
        return afq::check_launch("afq_trotter");                                                       \
    }
    AFQ_F64_RES(1)
    AFQ_F64_RES(2)
    AFQ_F64_RES(3)
#undef AFQ_F64_RES
    afq::set_error("afq_trotter: no resident kernel for o=%d nspin=%d", o, nspin);
    return -1;
}

template <typename R>
int trotter_entry(const char* name, int nw, int n, int o, int nspin, int ldv, int n_terms,
                  double sqrt_dt, const R* exp_h1_half, const R* vhs, const double* win_up,
                  const double* win_dn, double* wout_up, double* wout_dn, hipStream_t st) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "%s: bad sizes nw=%d n=%d o=%d", name, nw, n, o);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "%s: nspin must be 1 or 2", name);
    AFQ_REQUIRE(ldv >= n && ldv % 4 == 0, "%s: ldv=%d must be >= n and a multiple of 4", name, ldv);
    AFQ_REQUIRE(n_terms >= 0, "%s: n_terms < 0", name);
    AFQ_REQUIRE(exp_h1_half && vhs && win_up && wout_up, "%s: null pointer", name);
    AFQ_REQUIRE(nspin == 1 || (win_dn && wout_dn), "%s: null spin-down pointer", name);
    AFQ_REQUIRE(win_up != wout_up && (nspin == 1 || win_dn != wout_dn), "%s: output aliases input", name);
    const int NP = afq::round_up(n, 16);
    // column tiles per workgroup: two (groups of 32 columns) once o > 16
    const int nc = o > 16 ? 2 : 1;
    const int ntile = NP / 16;
    const int rpw = afq::ceil_div(ntile, 4);
    AFQ_REQUIRE(o <= 64, "%s: o=%d exceeds the supported maximum of 64", name, o);
    AFQ_REQUIRE(rpw <= 5, "%s: n=%d exceeds the supported maximum of 320", name, n);
    AFQ_REQUIRE((long long)nw * nspin * afq::ceil_div(o, 32) < (1LL << 31), "%s: too many walkers", name);
    AFQ_REQUIRE((size_t)2 * NP * nc * 16 * sizeof(R) <= 160 * 1024,
                "%s: n=%d o=%d needs more than 160 KB of LDS", name, n, o);
    const double2* wu = (const double2*)win_up;
    const double2* wd = (const double2*)(nspin == 2 ? win_dn : win_up);
    double2* ou = (double2*)wout_up;
    double2* od = (double2*)(nspin == 2 ? wout_dn : wout_up);
    if constexpr (sizeof(R) == sizeof(double)) {
        // VHS resident in registers + LDS up to 128 rows, both spins per workgroup
        if (NP <= F64_RES_MAX_NP) {
            int ngrp, chunk, nct;
            f64_resident_shape(o, nspin, ngrp, chunk, nct);
            if (f64_resident_lds(n, nct) <= 160 * 1024 && (long long)nw * ngrp < (1LL << 31))
                return launch_trotter_f64_resident(nw, n, o, nspin, ldv, n_terms, sqrt_dt, exp_h1_half,
                                                   vhs, wu, wd, ou, od, st);
        }
    }
    if constexpr (sizeof(R) == sizeof(float)) {
        // VHS resident on chip when one wave per row tile fits (<= 16 waves)
        // and the LDS share of V plus T fits in 160 KB
        if (ntile <= 16 && resident_lds(n) <= 160 * 1024)
            return launch_trotter_resident(nw, n, o, nspin, ldv, n_terms, sqrt_dt, exp_h1_half,
                                           vhs, wu, wd, ou, od, st);
    }
#define AFQ_TROTTER_CASE(NC_, RPW_)                                                            \
    if (nc == NC_ && rpw == RPW_)                                                              \
        return launch_trotter<R, NC_, RPW_>(nw, n, o, nspin, ldv, n_terms, sqrt_dt, exp_h1_half, \
                                            vhs, wu, wd, ou, od, st);
    AFQ_TROTTER_CASE(1, 1)
    AFQ_TROTTER_CASE(1, 2)
    AFQ_TROTTER_CASE(1, 3)
    AFQ_TROTTER_CASE(1, 4)
    AFQ_TROTTER_CASE(1, 5)
    AFQ_TROTTER_CASE(2, 1)
    AFQ_TROTTER_CASE(2, 2)
    AFQ_TROTTER_CASE(2, 3)
    AFQ_TROTTER_CASE(2, 4)
    AFQ_TROTTER_CASE(2, 5)
#undef AFQ_TROTTER_CASE
    afq::set_error("%s: no kernel for n=%d o=%d", name, n, o);
    return -1;
}

// ---------------------------------------------------------------------------
// Adjoint of the fp64 Trotter step (restricted walkers, n <= 128, o <= 32),
//   w' = E u,  u = sum_{k<=K} t_k,  t_0 = E w,  t_k = (a / k) V t_{k-1},
//   a = i sqrt(dt), V and E symmetric:
// given w'_bar, the reverse recursion
//   u_bar = E^T w'_bar,  t_bar_K = u_bar,
//   t_bar_{k-1} = u_bar + conj(a / k) V^H t_bar_k = u_bar + conj(V) tb_k,
//   tb_k = conj(a / k) t_bar_k,  w_bar = E^T t_bar_0,
//   V_bar = sum_k tb_k t_{k-1}^H
// (prop/chol_afqmc_ops.py:173-196, prop/utils.py:20-31 differentiated).  One
// workgroup per walker, one wave per 16-row tile, V resident on chip as in the
// forward resident kernel.  The forward terms t_0..t_{K-1}, the scaled
// adjoints tb_1..tb_K, u and t_bar_0 go to the caller's workspace ([2K + 2]
// [nw][n][16 NCT] complex, padding columns zero); V_bar = sum_k tb_k
// t_{k-1}^H and the walker's share of E's adjoint, Re(w'_bar u^H + t_bar_0
// w^H), are products of those, each wave one 16-row slab.
// ---------------------------------------------------------------------------
template <int NCT>
__global__ __launch_bounds__(512, 1) void trotter_adjoint_f64_kernel(
    int nw, int n, int o, int ldv, int n_terms, double sqrt_dt, const double* __restrict__ E,
    const double* __restrict__ Et, const double* __restrict__ vhs, long long vplane,
    const double2* __restrict__ w_in, const double2* __restrict__ wbar_out, double2* __restrict__ wbar_in,
    double* __restrict__ vbar, double* __restrict__ ebar, double2* __restrict__ ws) {
    constexpr int KRB = f64_res_krb(NCT);
    constexpr int TC = 16 * NCT;
    using MF = afq::mfma16<double>;
    using ACC = MF::acc;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int NP = afq::round_up(n, 16);
    const int nkb = NP / 16;
    const int nkl = nkb > KRB ? nkb - KRB : 0;
    double* Tr = reinterpret_cast<double*>(lds_raw);
    double* Ti = Tr + NP * TC;
    double* VL = Ti + NP * TC;  // [wave][nkl][4 kk][2 planes][64 lanes]

    const int w = blockIdx.x;
    const long long woff = (long long)w * n * o;
    const double* Vr = vhs + (long long)w * n * ldv;
    const double* Vi = Vr + vplane;
    // term k of this walker in a workspace array: (n, TC) complex
    const long long tstride = (long long)nw * n * TC;
    auto slot = [&](double2* base, int k) { return base + k * tstride + (long long)w * n * TC; };
    double2* terms = ws;                                  // t_0 .. t_{K-1}
    double2* tbars = ws + (long long)n_terms * tstride;   // tb_1 .. tb_K
    double2* u_ws = slot(ws, 2 * n_terms);                // u
    double2* y0_ws = slot(ws, 2 * n_terms + 1);           // t_bar_0

    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int lane = tid & 63;
    const int rt = tid >> 6;
    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    const int row = rt * 16 + lr;

    auto load4 = [&](const double* A, int kb, double (&x)[4]) {
        const int k0 = kb * 16 + lk;
        const bool ok = row < n && k0 < n;
        const double4 v = *reinterpret_cast<const double4*>(A + (long long)(ok ? row : 0) * ldv + (ok ? k0 : 0));
        x[0] = ok ? v.x : 0.0; x[1] = ok ? v.y : 0.0; x[2] = ok ? v.z : 0.0; x[3] = ok ? v.w : 0.0;
    };
    // T <- walker-shaped (n, o) complex array, zero padded to NP x TC
    auto load_T = [&](const double2* src) {
        for (int e = tid; e < NP * TC; e += nthreads) {
            const int p = e / TC;
            const int c = e - p * TC;
            const double2 x = (p < n && c < o) ? src[woff + (long long)p * o + c] : make_double2(0.0, 0.0);
            Tr[e] = x.x;
            Ti[e] = x.y;
        }
    };
    ACC P[NCT][2], out[NCT][2], cur[NCT][2];
    auto e_product = [&](const double* M) {
        double a[F64_RES_MAX_NP / 16][4];
#pragma unroll
        for (int kb = 0; kb < F64_RES_MAX_NP / 16; ++kb)
            if (kb < nkb) load4(M, kb, a[kb]);
#pragma unroll
        for (int c = 0; c < NCT; ++c) P[c][0] = P[c][1] = MF::zero();
#pragma unroll
        for (int kb = 0; kb < F64_RES_MAX_NP / 16; ++kb) {
            if (kb < nkb) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int k = kb * 16 + lk + kk;
#pragma unroll
                    for (int c = 0; c < NCT; ++c) {
                        P[c][0] = MF::fma(a[kb][kk], Tr[k * TC + c * 16 + lr], P[c][0]);
                        P[c][1] = MF::fma(a[kb][kk], Ti[k * TC + c * 16 + lr], P[c][1]);
                    }
                }
            }
        }
    };
    auto store_T = [&](const ACC (&X)[NCT][2]) {
#pragma unroll
        for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = rt * 16 + MF::row(lane, r);
                Tr[rr * TC + c * 16 + lr] = X[c][0][r];
                Ti[rr * TC + c * 16 + lr] = X[c][1][r];
            }
    };
    // accumulator rows < n of X (times s = sr + i si) to a (n, TC) workspace term
    auto store_term = [&](double2* dst, const ACC (&X)[NCT][2], double sr, double si) {
#pragma unroll
        for (int c = 0; c < NCT; ++c)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rr = rt * 16 + MF::row(lane, r);
                const double xr = X[c][0][r], xi = X[c][1][r];
                if (rr < n) dst[(long long)rr * TC + c * 16 + lr] = make_double2(sr * xr - si * xi, sr * xi + si * xr);
            }
    };
    // accumulator rows < n, columns < o of X to a walker-shaped output
    auto store_walker = [&](double2* dst, const ACC (&X)[NCT][2]) {
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const int col = c * 16 + lr;
            if (col < o) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rr = rt * 16 + MF::row(lane, r);
                    if (rr < n) dst[woff + (long long)rr * o + col] = make_double2(X[c][0][r], X[c][1][r]);
                }
            }
        }
    };
    // P += V T (conj: += conj(V) T) for one k with A = (ar, ai)
    auto vt_step = [&](int k, double ar, double ai) {
#pragma unroll
        for (int c = 0; c < NCT; ++c) {
            const double br = Tr[k * TC + c * 16 + lr];
            const double bi = Ti[k * TC + c * 16 + lr];
            P[c][0] = MF::fma(ar, br, P[c][0]);
            P[c][1] = MF::fma(ar, bi, P[c][1]);
            P[c][0] = MF::fma(ai, -bi, P[c][0]);
            P[c][1] = MF::fma(ai, br, P[c][1]);
        }
    };

    // ---- forward: t_0 = E w, then V resident ----
    load_T(w_in);
    __syncthreads();
    e_product(E);
    store_term(slot(terms, 0), P, 1.0, 0.0);
    double vr[KRB][4], vi[KRB][4];
#pragma unroll
    for (int kb = 0; kb < KRB; ++kb) {
        load4(Vr, kb < nkb ? kb : 0, vr[kb]);
        load4(Vi, kb < nkb ? kb : 0, vi[kb]);
    }
    for (int kb = KRB; kb < nkb; ++kb) {
        double xr[4], xi[4];
        load4(Vr, kb, xr);
        load4(Vi, kb, xi);
        double* dst = VL + ((size_t)(rt * nkl + (kb - KRB)) * 8) * 64 + lane;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            dst[(kk * 2 + 0) * 64] = xr[kk];
            dst[(kk * 2 + 1) * 64] = xi[kk];
        }
    }
    __syncthreads();
    store_T(P);
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
        out[c][0] = P[c][0];
        out[c][1] = P[c][1];
    }
    __syncthreads();

    // V (sgn = 1) or conj(V) (sgn = -1) times T into P
    auto v_product = [&](double sgn) {
#pragma unroll
        for (int c = 0; c < NCT; ++c) P[c][0] = P[c][1] = MF::zero();
#pragma unroll
        for (int kb = 0; kb < KRB; ++kb) {
            if (kb < nkb) {
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) vt_step(kb * 16 + lk + kk, vr[kb][kk], sgn * vi[kb][kk]);
            }
        }
        for (int kb = KRB; kb < nkb; ++kb) {
            const double* src = VL + ((size_t)(rt * nkl + (kb - KRB)) * 8) * 64 + lane;
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
                vt_step(kb * 16 + lk + kk, src[(kk * 2 + 0) * 64], sgn * src[(kk * 2 + 1) * 64]);
        }
    };

#pragma unroll 1
    for (int k = 1; k <= n_terms; ++k) {
        v_product(1.0);
        const double ck = sqrt_dt / (double)k;
#pragma unroll
        for (int c = 0; c < NCT; ++c) {  // t_k = i ck V t_{k-1}
            const ACC re = -ck * P[c][1];
            const ACC im = ck * P[c][0];
            P[c][0] = re;
            P[c][1] = im;
            out[c][0] += re;
            out[c][1] += im;
        }
        if (k < n_terms) store_term(slot(terms, k), P, 1.0, 0.0);
        __syncthreads();
        store_T(P);
        __syncthreads();
    }
    store_term(u_ws, out, 1.0, 0.0);

    // ---- reverse: u_bar = E^T w'_bar ----
    __syncthreads();
    load_T(wbar_out);
    __syncthreads();
    e_product(Et);
#pragma unroll
    for (int c = 0; c < NCT; ++c) {
        out[c][0] = P[c][0];  // u_bar
        out[c][1] = P[c][1];
        cur[c][0] = P[c][0];  // t_bar_K
        cur[c][1] = P[c][1];
    }
    __syncthreads();
    store_T(cur);
    __syncthreads();
#pragma unroll 1
    for (int k = n_terms; k >= 1; --k) {
        const double ck = sqrt_dt / (double)k;
        // tb_k = conj(i ck) t_bar_k = -i ck t_bar_k
        store_term(slot(tbars, k - 1), cur, 0.0, -ck);
        v_product(-1.0);
#pragma unroll
        for (int c = 0; c < NCT; ++c) {  // t_bar_{k-1} = u_bar - i ck conj(V) t_bar_k
            cur[c][0] = out[c][0] + ck * P[c][1];
            cur[c][1] = out[c][1] - ck * P[c][0];
        }
        __syncthreads();
        store_T(cur);
        __syncthreads();
    }
    store_term(y0_ws, cur, 1.0, 0.0);
    e_product(Et);
    store_walker(wbar_in, P);

    // ---- V_bar = sum_k tb_k t_{k-1}^H: this wave's 16 rows, all columns ----
    __syncthreads();  // every wave's workspace terms are written
    {
        constexpr int NTM = F64_RES_MAX_NP / 16;
        ACC A[NTM][2];
#pragma unroll
        for (int ct = 0; ct < NTM; ++ct) A[ct][0] = A[ct][1] = MF::zero();
        const double2* tb0 = slot(tbars, 0);
        const double2* t0 = slot(terms, 0);
        const int nj = n_terms * TC / 16;  // 16-deep blocks of the (term, column) index
        const int rowa = row < n ? row : 0;
        for (int jb = 0; jb < nj; ++jb) {
            const int term = (jb * 16) / TC;
            const int c0 = jb * 16 - term * TC + lk;
            double2 a[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const double2 v = tb0[term * tstride + (long long)rowa * TC + c0 + kk];
                a[kk] = row < n ? v : make_double2(0.0, 0.0);
            }
#pragma unroll
            for (int ct = 0; ct < NTM; ++ct) {
                if (ct < nkb) {
                    const int q = ct * 16 + lr;
                    const int qa = q < n ? q : 0;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        double2 b = t0[term * tstride + (long long)qa * TC + c0 + kk];
                        if (q >= n) b = make_double2(0.0, 0.0);
                        // A . conj(B^T): B' = (br, -bi)
                        A[ct][0] = MF::fma(a[kk].x, b.x, A[ct][0]);
                        A[ct][1] = MF::fma(a[kk].x, -b.y, A[ct][1]);
                        A[ct][0] = MF::fma(-a[kk].y, -b.y, A[ct][0]);
                        A[ct][1] = MF::fma(a[kk].y, b.x, A[ct][1]);
                    }
                }
            }
        }
        double* vbr = vbar + (long long)w * n * n;
        double* vbi = vbr + (long long)nw * n * n;
#pragma unroll
        for (int ct = 0; ct < NTM; ++ct) {
            const int q = ct * 16 + lr;
            if (ct < nkb && q < n) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int p = rt * 16 + MF::row(lane, r);
                    if (p < n) {
                        vbr[(long long)p * n + q] = A[ct][0][r];
                        vbi[(long long)p * n + q] = A[ct][1][r];
                    }
                }
            }
        }
    }

    // ---- E's adjoint, this walker's share: Re(w'_bar u^H + t_bar_0 w^H) ----
    {
        constexpr int NTM = F64_RES_MAX_NP / 16;
        ACC A[NTM];
#pragma unroll
        for (int ct = 0; ct < NTM; ++ct) A[ct] = MF::zero();
        const int rowa = row < n ? row : 0;
        // walker-shaped (n, o) or workspace (n, TC) element (p, c), zero outside
        auto wk = [&](const double2* src, int p, int c) {
            const bool ok = p < n && c < o;
            const double2 v = src[woff + (long long)(ok ? p : 0) * o + (ok ? c : 0)];
            return ok ? v : make_double2(0.0, 0.0);
        };
        auto wsv = [&](const double2* src, int p, int c) {
            const double2 v = src[(long long)(p < n ? p : 0) * TC + c];
            return p < n ? v : make_double2(0.0, 0.0);
        };
        for (int cb = 0; cb < NCT; ++cb) {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int c = cb * 16 + lk + kk;
                const double2 a1 = wk(wbar_out, rowa, c);
                const double2 a2 = wsv(y0_ws, row, c);
#pragma unroll
                for (int ct = 0; ct < NTM; ++ct) {
                    if (ct < nkb) {
                        const int q = ct * 16 + lr;
                        const double2 b1 = wsv(u_ws, q, c);
                        const double2 b2 = wk(w_in, q, c);
                        A[ct] = MF::fma(row < n ? a1.x : 0.0, b1.x, A[ct]);
                        A[ct] = MF::fma(row < n ? a1.y : 0.0, b1.y, A[ct]);
                        A[ct] = MF::fma(a2.x, b2.x, A[ct]);
                        A[ct] = MF::fma(a2.y, b2.y, A[ct]);
                    }
                }
            }
        }
        double* eb = ebar + (long long)w * n * n;
#pragma unroll
        for (int ct = 0; ct < NTM; ++ct) {
            const int q = ct * 16 + lr;
            if (ct < nkb && q < n) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int p = rt * 16 + MF::row(lane, r);
                    if (p < n) eb[(long long)p * n + q] = A[ct][r];
                }
            }
        }
    }
}

}  // namespace

extern "C" int afq_trotter(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                           const double* exp_h1_half, const double* vhs, const double* win_up,
                           const double* win_dn, double* wout_up, double* wout_dn, void* stream) {
    return trotter_entry<double>("afq_trotter", nw, n, o, nspin, ldv, n_terms, sqrt_dt, exp_h1_half,
                                 vhs, win_up, win_dn, wout_up, wout_dn, (hipStream_t)stream);
}

extern "C" int afq_trotter_f32(int nw, int n, int o, int nspin, int ldv, int n_terms,
                               double sqrt_dt, const float* exp_h1_half, const float* vhs,
                               const double* win_up, const double* win_dn, double* wout_up,
                               double* wout_dn, void* stream) {
    return trotter_entry<float>("afq_trotter_f32", nw, n, o, nspin, ldv, n_terms, sqrt_dt,
                                exp_h1_half, vhs, win_up, win_dn, wout_up, wout_dn,
                                (hipStream_t)stream);
}

extern "C" int afq_trotter_f32_grouped(int count, const afq_system_step* sys, int total_blocks,
                                       int max_n, int n_terms, void* stream) {
    AFQ_REQUIRE(count > 0 && total_blocks > 0, "afq_trotter_f32_grouped: count=%d total_blocks=%d",
                count, total_blocks);
    AFQ_REQUIRE(sys, "afq_trotter_f32_grouped: null pointer");
    AFQ_REQUIRE(n_terms >= 0, "afq_trotter_f32_grouped: n_terms < 0");
    AFQ_REQUIRE(max_n > 0 && max_n <= 256 && resident_lds(max_n) <= 160 * 1024,
                "afq_trotter_f32_grouped: max_n=%d needs more than 160 KB of LDS", max_n);
    const int NP = afq::round_up(max_n, 16);
    const size_t lds = resident_lds(max_n);
    if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_resident_grouped_kernel<RESIDENT_KRB>);
    hipLaunchKernelGGL((trotter_resident_grouped_kernel<RESIDENT_KRB>), dim3((unsigned)total_blocks),
                       dim3(64 * (NP / 16)), lds, (hipStream_t)stream, count, total_blocks, NP, n_terms,
                       sys);
    return afq::check_launch("afq_trotter_f32_grouped");
}

extern "C" int afq_trotter_adjoint(int nw, int n, int o, int ldv, int n_terms, double sqrt_dt,
                                   const double* exp_h1_half, const double* exp_h1_half_t, const double* vhs,
                                   const double* w_in, const double* w_bar_out, double* w_bar_in, double* vhs_bar,
                                   double* e_half_bar, double* workspace, void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "afq_trotter_adjoint: bad sizes nw=%d n=%d o=%d", nw, n, o);
    AFQ_REQUIRE(n <= F64_RES_MAX_NP && o <= 32, "afq_trotter_adjoint: n=%d o=%d beyond n <= %d, o <= 32", n, o,
                F64_RES_MAX_NP);
    AFQ_REQUIRE(ldv >= n && ldv % 4 == 0, "afq_trotter_adjoint: ldv=%d must be >= n and a multiple of 4", ldv);
    AFQ_REQUIRE(n_terms >= 1, "afq_trotter_adjoint: n_terms=%d < 1", n_terms);
    AFQ_REQUIRE(exp_h1_half && exp_h1_half_t && vhs && w_in && w_bar_out && w_bar_in && vhs_bar && e_half_bar &&
                    workspace,
                "afq_trotter_adjoint: null pointer");
    const int nct = afq::ceil_div(o, 16);
    const int NP = afq::round_up(n, 16);
    const long long vplane = (long long)nw * n * ldv;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((unsigned)nw), block(64 * (NP / 16));
#define AFQ_TROTTER_ADJ(NCT_)                                                                                    \
    if (nct == NCT_) {                                                                                           \
        const size_t lds = f64_resident_lds(n, NCT_);                                                            \
        AFQ_REQUIRE(lds <= 160 * 1024, "afq_trotter_adjoint: n=%d needs %zu bytes of LDS", n, lds);              \
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_adjoint_f64_kernel<NCT_>);                \
        hipLaunchKernelGGL((trotter_adjoint_f64_kernel<NCT_>), grid, block, lds, st, nw, n, o, ldv, n_terms, sqrt_dt, \
                           exp_h1_half, exp_h1_half_t, vhs, vplane, (const double2*)w_in, (const double2*)w_bar_out, \
                           (double2*)w_bar_in, vhs_bar, e_half_bar, (double2*)workspace);                        \
        return afq::check_launch("afq_trotter_adjoint");                                                        \
    }
    AFQ_TROTTER_ADJ(1)
    AFQ_TROTTER_ADJ(2)
#undef AFQ_TROTTER_ADJ
    afq::set_error("afq_trotter_adjoint: no kernel for o=%d", o);
    return -1;
}
