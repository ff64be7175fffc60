// Split-bf16 ("bf16x3") Trotter step of the mixed-precision walker step.
//
//   W <- E . Taylor_m(i sqrt(dt) V) . E . W,  E = exp(-dt h1 / 2)
//
// (reference: prop/chol_afqmc_ops.py:147-166 and :115-128).  Inputs and
// outputs are those of the fp32 step: fp32 E (n x ldv) and VHS planes, fp64
// complex walkers.  Arithmetic:
//
// * Every matrix product runs on v_mfma_f32_16x16x32_bf16 with both operands
//   split as x = hi + lo (hi = bf16(x), lo = bf16(x - hi), ~16 significant
//   bits) and accumulated in fp32 as al.bh + ah.bl + ah.bh.
// * The Taylor products V.T all carry the factor sqrt(dt) / k, so the split
//   error (~2^-17 relative) enters the walker scaled by sqrt(dt) ||V||.
// * E.X is formed as X + (E - I).X: X is added exactly in fp32 (it is the
//   accumulator's initial value) and only (E - I), of norm ~dt ||h1|| / 2,
//   goes through the split product.
// * The running sum and every Taylor term are fp32 accumulators.
//
// Layout: one workgroup per (walker, spin, 16-column tile of the walker),
// ceil(ntile / 2) waves, each owning two 16-row tiles of V.  The walker's V
// is resident on chip for all Taylor terms: k-blocks (32 wide) 0..KR-1 as
// split MFMA A fragments in registers (160 VGPRs per lane), the remaining NL
// k-blocks in LDS in fragment order (lane-contiguous 16-byte slots).  The
// current Taylor term T (n x 16) sits in LDS as four bf16 planes (re hi/lo,
// im hi/lo), column-major so that each lane's B fragment (8 consecutive k of
// one column) is one 16-byte read.  The workgroups of one walker run back to
// back on one XCD and share V through L2.
#include <cstdint>

#include "afqmc_hip.h"
#include "common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned v4u32 __attribute__((ext_vector_type(4)));

constexpr int RT = 2;  // 16-row tiles per wave
constexpr int KR = 5;  // 32-wide k-blocks of V held in registers
constexpr int NL_MAX = 2;

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 neg8(bf16x8 x) {
    uint4 u = __builtin_bit_cast(uint4, x);
    u.x ^= 0x80008000u;
    u.y ^= 0x80008000u;
    u.z ^= 0x80008000u;
    u.w ^= 0x80008000u;
    return __builtin_bit_cast(bf16x8, u);
}

struct split8 {
    bf16x8 hi, lo;
};

__device__ __forceinline__ split8 split_f32x8(const float (&x)[8]) {
    split8 s;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const __bf16 h = (__bf16)x[j];
        s.hi[j] = h;
        s.lo[j] = (__bf16)(x[j] - (float)h);
    }
    return s;
}

// Eight bf16_pair words (hi in the low half, lo in the high half), loaded as
// raw 32-bit values, regrouped into the hi and lo operand vectors: one byte
// permute per pair of elements and half.
__device__ __forceinline__ split8 unpack_pairs(const float (&x)[8]) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t a = __builtin_bit_cast(uint32_t, x[2 * j]);
        const uint32_t b = __builtin_bit_cast(uint32_t, x[2 * j + 1]);
        h[j] = __builtin_amdgcn_perm(b, a, 0x05040100u);
        l[j] = __builtin_amdgcn_perm(b, a, 0x07060302u);
    }
    split8 s;
    s.hi = __builtin_bit_cast(bf16x8, make_uint4(h[0], h[1], h[2], h[3]));
    s.lo = __builtin_bit_cast(bf16x8, make_uint4(l[0], l[1], l[2], l[3]));
    return s;
}

// Raw buffer loads: a 32-bit byte offset from a per-array base in SGPRs, and
// zeros for offsets at or past the array's size (gfx9 descriptor word 3).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr unsigned kOutside = 0x80000000u;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float4 load_f4(rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// 16 bytes per lane from buffer r at byte offset off straight into LDS at
// slot + lane (slot wave-uniform); zeros for offsets past the buffer.  The data
// is there once the issuing wave has passed s_waitcnt vmcnt(0).
__device__ __forceinline__ void load_lds16(rsrc_t r, uint4* slot, unsigned off) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_ptr_t)slot, 16, off, 0, 0, 0);
}

__device__ __forceinline__ double2 load_d2(rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}

// Eight consecutive k (k0 .. k0+7, k0 a multiple of 8) of row `row` of an
// n x n fp32 matrix stored with leading dimension ld (a multiple of 4) in the
// buffer r of n * ld floats: rows >= n read as zero (outside the buffer), k
// blocks starting at or past n are skipped; columns n..ld-1 inside a row are
// read as stored.  Nothing here waits on the loaded data.
__device__ __forceinline__ void load_row8(rsrc_t r, int ld, int n, int row, int k0, float (&x)[8]) {
    const unsigned base = (unsigned)(row * ld + k0) * 4u;
    const float4 a = load_f4(r, k0 < n ? base : kOutside);
    const float4 c = load_f4(r, k0 + 4 < n ? base + 16u : kOutside);
    x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
    x[4] = c.x; x[5] = c.y; x[6] = c.z; x[7] = c.w;
}

__device__ __forceinline__ float load_f1(rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// The A fragment V[16 bi + lr][32 kb + 8 q .. + 7] of a symmetric V stored as
// its upper 16 x 16 blocks (block (r, c), r <= c, numbered row by row, entry
// (pi, qi) at 16 pi + qi; afq_vhs_sym's column order): a row of an upper
// block (two 16-byte loads), or, below the diagonal, a column of the stored
// transposed block (eight 4-byte loads 64 bytes apart, each coalesced over the
// 16 lanes of a column); zero past the blocks.  Nothing here waits on the data.
__device__ __forceinline__ void load_frag_sym(rsrc_t r, int nb, int bi, int kb, int lr, int q, float (&x)[8]) {
    const int bj = 2 * kb + (q >> 1);
    const int cb = 8 * (q & 1);
    if (bi > bj) {
        const int blk = bj * nb - ((bj * (bj - 1)) >> 1) + (bi - bj);
        const unsigned base = (bi < nb) ? (unsigned)(blk * 256 + cb * 16 + lr) * 4u : kOutside;
#pragma unroll
        for (int j = 0; j < 8; ++j) x[j] = load_f1(r, base + 64u * j);
    } else {
        const int blk = bi * nb - ((bi * (bi - 1)) >> 1) + (bj - bi);
        const unsigned base = (bj < nb) ? (unsigned)(blk * 256 + lr * 16 + cb) * 4u : kOutside;
        const float4 a = load_f4(r, base);
        const float4 c = load_f4(r, base + 16u);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        x[4] = c.x; x[5] = c.y; x[6] = c.z; x[7] = c.w;
    }
}

// x through an empty asm: the compiler recomputes what depends on it instead
// of keeping (and spilling) values from before a long loop
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

// split of x - I restricted to row `row` (< n), columns k0..k0+7; branch-free
__device__ __forceinline__ split8 split_minus_identity(const float (&x)[8], int row, int n, int k0) {
    float y[8];
    const int d = row < n ? row - k0 : -1;
#pragma unroll
    for (int j = 0; j < 8; ++j) y[j] = x[j] - (d == j ? 1.f : 0.f);
    return split_f32x8(y);
}

// SYMV: vhs holds each walker's V, and E holds exp(-dt h1 / 2) - I, as upper
// 16 x 16 blocks (afq_vhs_sym's column order, nb (nb + 1) / 2 * 256 entries per
// matrix) instead of (n, ldv) rows; both are symmetric, and every entry is
// already split: a bf16_pair word (the split the kernel otherwise does on
// load, from the same fp32 value), so the fragments are regrouped, not converted.
// GROUPED: the workgroup's system is found in the descriptor table S by the
// prefix sums tblock0 (count systems, restricted walkers, (n, ldv) layouts);
// its sizes and pointers replace the arguments, and blockIdx.x counts from
// the system's first workgroup.
template <int NL, bool SYMV = false, bool GROUPED = false>
__global__ __launch_bounds__(512, 1) void trotter_bf16x3_kernel(
    int nw, int n, int o, int nspin, int nct, int ldv, int n_terms, double sqrt_dt,
    const float* __restrict__ E, const float* __restrict__ vhs, long long vplane,
    const double2* __restrict__ win_up, const double2* __restrict__ win_dn,
    double2* __restrict__ wout_up, double2* __restrict__ wout_dn, int group_map, int count = 0,
    const afq_system_step* __restrict__ S = nullptr, const float* __restrict__ Elast = nullptr,
    int final_half = 1, int first_full = 0) {
    constexpr int KMAX = KR + NL;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    int b = blockIdx.x;
    if constexpr (GROUPED) {
        const int t = afq::xcd_contiguous(blockIdx.x, gridDim.x);
        const int si = afq::find_group(count, t, [&](int i) { return S[i].tblock0; });
        const afq_system_step& d = S[si];
        n = d.n;
        o = d.o;
        ldv = d.ldv;
        nw = d.nw;
        nspin = 1;
        nct = afq::ceil_div(o, 16);
        sqrt_dt = d.sqrt_dt;
        E = first_full ? d.exp_h1 : d.exp_h1_half;
        Elast = d.exp_h1_half;
        vhs = d.vhs;
        if constexpr (SYMV) {
            const int nbg = afq::ceil_div(n, 16);
            vplane = (long long)nw * nbg * (nbg + 1) / 2 * 256;
        } else {
            vplane = (long long)nw * n * ldv;
        }
        win_up = win_dn = (const double2*)d.walkers_in;
        wout_up = wout_dn = (double2*)d.walkers_out;
        group_map = 0;
        b = t - d.tblock0;
    }
    const int NP = afq::round_up(n, 16);
    const int nk = afq::ceil_div(NP, 32);  // == KMAX when NL > 0, <= KR otherwise
    const int TP = nk * 32 + 8;            // bf16 pitch of one column of T
    __bf16* Tt = reinterpret_cast<__bf16*>(lds_raw);  // [4 planes][16 cols][TP]
    uint4* VL = reinterpret_cast<uint4*>(lds_raw + (size_t)4 * 16 * TP * sizeof(__bf16));

    const int G = nspin * nct;
    int w, sub;
    if (group_map) {
        w = (b / (8 * G)) * 8 + (b % 8);
        sub = (b / 8) % G;
    } else {
        w = b / G;
        sub = b % G;
    }
    const int s = sub / nct;
    const int c0 = (sub - s * nct) * 16;
    const long long woff = (long long)w * n * o;
    double2* Wout = (s == 0 ? wout_up : wout_dn) + woff;
    const rsrc_t Wr = make_rsrc((s == 0 ? win_up : win_dn) + woff, (unsigned)n * o * 16u);
    const unsigned mat_bytes = (unsigned)n * ldv * 4u;
    const int nbv = NP / 16;
    const long long vrow = SYMV ? (long long)nbv * (nbv + 1) / 2 * 256 : (long long)n * ldv;
    const rsrc_t Er = make_rsrc(E, SYMV ? (unsigned)vrow * 4u : mat_bytes);
    const rsrc_t ErL = make_rsrc(Elast ? Elast : E, SYMV ? (unsigned)vrow * 4u : mat_bytes);
    const rsrc_t Vr = make_rsrc(vhs + (long long)w * vrow, (unsigned)vrow * 4u);
    const rsrc_t Vi = make_rsrc(vhs + vplane + (long long)w * vrow, (unsigned)vrow * 4u);

    // blockDim.x = 64 * nk: one wave per pair of 16-row tiles
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lr = lane & 15;
    const int q = lane >> 4;
    const int kq = 8 * q;
    const int col = c0 + lr;
    const int row0 = wave * RT * 16 + lr;  // A-fragment row of tile 0 (tile t: + 16 t)

    auto read_b = [&](int kb, int plane) {
        return *reinterpret_cast<const bf16x8*>(Tt + (plane * 16 + lr) * TP + kb * 32 + kq);
    };
    // W[p][j], zero outside n x o (rows p >= n fall outside the buffer)
    auto load_w = [&](int p, int j) {
        return load_d2(Wr, j < o ? (unsigned)(p * o + j) * 16u : kOutside);
    };
    // all fragments of this wave's rows of the one-body matrix in src, as raw
    // fp32 (one round trip) (SYMV: given as its upper blocks too, read like V)
    auto load_e = [&](const rsrc_t& src, float (&er)[RT][KMAX][8], int r0, int k8) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int kb = 0; kb < KMAX; ++kb) {
                if constexpr (SYMV)
                    load_frag_sym(src, nbv, (r0 >> 4) + t, kb, r0 & 15, k8 >> 3, er[t][kb]);
                else
                    load_row8(src, ldv, n, r0 + 16 * t, kb * 32 + k8, er[t][kb]);
            }
    };
    // acc += (E - I) . T
    auto e_product = [&](f32x4 (&acc)[RT][2], const float (&er)[RT][KMAX][8], int r0, int k8) {
#pragma unroll
        for (int kb = 0; kb < KMAX; ++kb) {
            if (kb < nk) {
                split8 ef[RT];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    if constexpr (SYMV)
                        ef[t] = unpack_pairs(er[t][kb]);
                    else
                        ef[t] = split_minus_identity(er[t][kb], r0 + 16 * t, n, kb * 32 + k8);
                }
                const bf16x8 brh = read_b(kb, 0), brl = read_b(kb, 1);
                const bf16x8 bih = read_b(kb, 2), bil = read_b(kb, 3);
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][0] = mfma_bf16(ef[t].lo, brh, acc[t][0]);
                    acc[t][1] = mfma_bf16(ef[t].lo, bih, acc[t][1]);
                }
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][0] = mfma_bf16(ef[t].hi, brl, acc[t][0]);
                    acc[t][1] = mfma_bf16(ef[t].hi, bil, acc[t][1]);
                }
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][0] = mfma_bf16(ef[t].hi, brh, acc[t][0]);
                    acc[t][1] = mfma_bf16(ef[t].hi, bih, acc[t][1]);
                }
            }
        }
    };
    // T <- split(X): lane writes rows 4q..4q+3 of its tiles in column lr
    auto store_t = [&](const f32x4 (&X)[RT][2]) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                bf16x8 hl;  // [0..3] hi, [4..7] lo
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const __bf16 h = (__bf16)X[t][c][r];
                    hl[r] = h;
                    hl[4 + r] = (__bf16)(X[t][c][r] - (float)h);
                }
                const uint4 u = __builtin_bit_cast(uint4, hl);
                const int k = (wave * RT + t) * 16 + 4 * q;
                *reinterpret_cast<uint2*>(Tt + ((2 * c + 0) * 16 + lr) * TP + k) = make_uint2(u.x, u.y);
                *reinterpret_cast<uint2*>(Tt + ((2 * c + 1) * 16 + lr) * TP + k) = make_uint2(u.z, u.w);
            }
    };

    // ---- prologue: walker columns, accumulator rows and E in flight together ----
    f32x4 acc[RT][2];
    {
        // accumulator element r of tile t: row (wave*RT + t)*16 + 4q + r, column
        // col; the waves' accumulators cover all nk * 32 rows of T
        double2 wa[RT][4];
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                wa[t][r] = load_w((wave * RT + t) * 16 + 4 * q + r, col);
            }
        float er[RT][KMAX][8];
        load_e(Er, er, row0, kq);
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                acc[t][0][r] = (float)wa[t][r].x;
                acc[t][1][r] = (float)wa[t][r].y;
            }
        store_t(acc);  // T = W
        __syncthreads();
        // ---- T0 = E W ----
        e_product(acc, er, row0, kq);
    }
    f32x4 out[RT][2];
#pragma unroll
    for (int t = 0; t < RT; ++t) {
        out[t][0] = acc[t][0];
        out[t][1] = acc[t][1];
    }
    __syncthreads();
    store_t(acc);

    // ---- V resident: the LDS k-blocks first (one round trip, registers
    // released after the LDS store), then the register k-blocks ----
    if constexpr (NL > 0) {
        float xl[RT][NL][2][8];
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                const int k0 = (KR + l) * 32 + kq;
                if constexpr (SYMV) {
                    load_frag_sym(Vr, nbv, wave * RT + t, KR + l, lr, q, xl[t][l][0]);
                    load_frag_sym(Vi, nbv, wave * RT + t, KR + l, lr, q, xl[t][l][1]);
                } else {
                    load_row8(Vr, ldv, n, row0 + 16 * t, k0, xl[t][l][0]);
                    load_row8(Vi, ldv, n, row0 + 16 * t, k0, xl[t][l][1]);
                }
            }
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                uint4* dst = VL + ((size_t)((wave * RT + t) * NL + l) * 4) * 64 + lane;
                const split8 a = SYMV ? unpack_pairs(xl[t][l][0]) : split_f32x8(xl[t][l][0]);
                const split8 c = SYMV ? unpack_pairs(xl[t][l][1]) : split_f32x8(xl[t][l][1]);
                dst[0 * 64] = __builtin_bit_cast(uint4, a.hi);
                dst[1 * 64] = __builtin_bit_cast(uint4, a.lo);
                dst[2 * 64] = __builtin_bit_cast(uint4, c.hi);
                dst[3 * 64] = __builtin_bit_cast(uint4, c.lo);
            }
    }
    split8 vr[RT][KR], vi[RT][KR];
    {
        float xv[RT][KR][2][8];
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int kb = 0; kb < KR; ++kb) {
                const int k0 = kb * 32 + kq;
                if constexpr (SYMV) {
                    load_frag_sym(Vr, nbv, wave * RT + t, kb, lr, q, xv[t][kb][0]);
                    load_frag_sym(Vi, nbv, wave * RT + t, kb, lr, q, xv[t][kb][1]);
                } else {
                    load_row8(Vr, ldv, n, row0 + 16 * t, k0, xv[t][kb][0]);
                    load_row8(Vi, ldv, n, row0 + 16 * t, k0, xv[t][kb][1]);
                }
            }
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int kb = 0; kb < KR; ++kb) {
                if constexpr (SYMV) {
                    vr[t][kb] = unpack_pairs(xv[t][kb][0]);
                    vi[t][kb] = unpack_pairs(xv[t][kb][1]);
                } else {
                    vr[t][kb] = split_f32x8(xv[t][kb][0]);
                    vi[t][kb] = split_f32x8(xv[t][kb][1]);
                }
            }
    }
    __syncthreads();

    // P += V . T for one k-block, A fragments (ar, ai) per row tile
    auto vt_block = [&](f32x4 (&P)[RT][2], int kb, const split8 (&ar)[RT], const split8 (&ai)[RT]) {
        const bf16x8 brh = read_b(kb, 0), brl = read_b(kb, 1);
        bf16x8 bih = read_b(kb, 2), bil = read_b(kb, 3);
        // re += Vr Tr, im += Vr Ti
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][0] = mfma_bf16(ar[t].lo, brh, P[t][0]);
            P[t][1] = mfma_bf16(ar[t].lo, bih, P[t][1]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][0] = mfma_bf16(ar[t].hi, brl, P[t][0]);
            P[t][1] = mfma_bf16(ar[t].hi, bil, P[t][1]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][0] = mfma_bf16(ar[t].hi, brh, P[t][0]);
            P[t][1] = mfma_bf16(ar[t].hi, bih, P[t][1]);
        }
        // im += Vi Tr, re += Vi (-Ti)
        bih = neg8(bih);
        bil = neg8(bil);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][1] = mfma_bf16(ai[t].lo, brh, P[t][1]);
            P[t][0] = mfma_bf16(ai[t].lo, bih, P[t][0]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][1] = mfma_bf16(ai[t].hi, brl, P[t][1]);
            P[t][0] = mfma_bf16(ai[t].hi, bil, P[t][0]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][1] = mfma_bf16(ai[t].hi, brh, P[t][1]);
            P[t][0] = mfma_bf16(ai[t].hi, bih, P[t][0]);
        }
    };

    // ---- Taylor terms: T_k = (i sqrt(dt) / k) V T_{k-1}, out += T_k ----
#pragma unroll 1
    for (int k = 1; k <= n_terms; ++k) {
        f32x4 P[RT][2];
#pragma unroll
        for (int t = 0; t < RT; ++t) P[t][0] = P[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KR; ++kb) {
            if (kb < nk) {
                split8 ar[RT], ai[RT];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    ar[t] = vr[t][kb];
                    ai[t] = vi[t][kb];
                }
                vt_block(P, kb, ar, ai);
            }
        }
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            split8 ar[RT], ai[RT];
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                const uint4* src = VL + ((size_t)((wave * RT + t) * NL + l) * 4) * 64 + lane;
                ar[t].hi = __builtin_bit_cast(bf16x8, src[0 * 64]);
                ar[t].lo = __builtin_bit_cast(bf16x8, src[1 * 64]);
                ai[t].hi = __builtin_bit_cast(bf16x8, src[2 * 64]);
                ai[t].lo = __builtin_bit_cast(bf16x8, src[3 * 64]);
            }
            vt_block(P, KR + l, ar, ai);
        }
        const float ck = (float)(sqrt_dt / (double)k);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            const f32x4 re = -ck * P[t][1];
            const f32x4 im = ck * P[t][0];
            P[t][0] = re;
            P[t][1] = im;
            out[t][0] += re;
            out[t][1] += im;
        }
        __syncthreads();
        store_t(P);
        __syncthreads();
    }

    // ---- W_out = E' out = out + (E' - I) out (E' = Elast, or E), or, with
    // final_half == 0, W_out = out: the step's last one-body half is left to
    // the next step, which applies it together with its own first half ----
    // (indices recomputed from an opaque lane id: nothing address-like stays
    // live across the Taylor loop)
    const int lane2 = opaque(lane);
    const int wave2 = opaque(wave);
    if (final_half) {
        float er[RT][KMAX][8];
        load_e(ErL, er, wave2 * RT * 16 + (lane2 & 15), 8 * (lane2 >> 4));
        __syncthreads();
        store_t(out);
        __syncthreads();
        e_product(out, er, wave2 * RT * 16 + (lane2 & 15), 8 * (lane2 >> 4));
    }
    const int col2 = c0 + (lane2 & 15);
    if (col2 < o) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = (wave2 * RT + t) * 16 + 4 * (lane2 >> 4) + r;
                if (row < n) Wout[row * o + col2] = make_double2((double)out[t][0][r], (double)out[t][1][r]);
            }
    }
}

// V and E - I as upper blocks, walkers with two items (nspin * ceil(o / 16) ==
// 2) and at least as many walkers as CUs: one persistent workgroup per CU loops over walkers
// w = blockIdx.x, + gridDim.x, ...; per walker
//   E W of both items (the second parked in its output rows) -> V into the
//   CU once -> the two Taylor series (the first sum parked in its output rows)
//   -> E - I -> both final one-body halves, while the next walker's columns
//   load;
// E - I stays in registers from one walker's final halves into the next
// walker's first ones, so it is loaded once per walker (the per-item launch
// loads it twice per item).  Parked values are fp32 sums stored as fp64, so
// reading them back is exact, and every product runs in the per-item kernel's
// order: the results equal trotter_bf16x3_kernel's bit for bit.
// MODE 1: both halves on E - I (as above).  MODE 0: no final halves (the sums
// are the output; the next step applies this step's last half together with
// its own first).  MODE 2: the final halves on Elast - I, and E - I loaded
// again at the top of every walker (nothing carried between walkers).
template <int NL, int MODE>
__global__ __launch_bounds__(512, 1) void trotter_bf16x3_pair_kernel(
    int nw, int n, int o, int nspin, int nct, int n_terms, double sqrt_dt, const float* __restrict__ E,
    const float* __restrict__ Elast, const float* __restrict__ vhs, long long vplane, const double2* __restrict__ win_up,
    const double2* __restrict__ win_dn, double2* __restrict__ wout_up, double2* __restrict__ wout_dn) {
    constexpr int KMAX = KR + NL;
    constexpr bool SYMV = true;
    const int ldv = 0;  // V and E - I are upper blocks
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int NP = afq::round_up(n, 16);
    const int nk = afq::ceil_div(NP, 32);
    const int TP = nk * 32 + 8;
    __bf16* Tt = reinterpret_cast<__bf16*>(lds_raw);  // [4 planes][16 cols][TP]
    uint4* VL = reinterpret_cast<uint4*>(lds_raw + (size_t)4 * 16 * TP * sizeof(__bf16));

    // the two items: spin s_i, first column c_i
    const int s1 = 1 / nct;
    const int c1 = (1 - s1 * nct) * 16;
    const unsigned mat_bytes = (unsigned)n * ldv * 4u;
    const int nbv = NP / 16;
    const long long vrow = SYMV ? (long long)nbv * (nbv + 1) / 2 * 256 : (long long)n * ldv;
    const rsrc_t Er = make_rsrc(E, SYMV ? (unsigned)vrow * 4u : mat_bytes);
    const rsrc_t ErL = make_rsrc(MODE == 2 ? Elast : E, SYMV ? (unsigned)vrow * 4u : mat_bytes);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int lr = lane & 15;
    const int q = lane >> 4;
    const int kq = 8 * q;
    const int row0 = wave * RT * 16 + lr;

    auto read_b = [&](int kb, int plane) {
        return *reinterpret_cast<const bf16x8*>(Tt + (plane * 16 + lr) * TP + kb * 32 + kq);
    };
    auto load_e = [&](const rsrc_t& src, float (&er)[RT][KMAX][8], int r0, int k8) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int kb = 0; kb < KMAX; ++kb) {
                if constexpr (SYMV)
                    load_frag_sym(src, nbv, (r0 >> 4) + t, kb, r0 & 15, k8 >> 3, er[t][kb]);
                else
                    load_row8(src, ldv, n, r0 + 16 * t, kb * 32 + k8, er[t][kb]);
            }
    };
    // acc += (E - I) . T
    auto e_product = [&](f32x4 (&acc)[RT][2], const float (&er)[RT][KMAX][8], int r0, int k8) {
#pragma unroll
        for (int kb = 0; kb < KMAX; ++kb) {
            if (kb < nk) {
                split8 ef[RT];
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    if constexpr (SYMV)
                        ef[t] = unpack_pairs(er[t][kb]);
                    else
                        ef[t] = split_minus_identity(er[t][kb], r0 + 16 * t, n, kb * 32 + k8);
                }
                const bf16x8 brh = read_b(kb, 0), brl = read_b(kb, 1);
                const bf16x8 bih = read_b(kb, 2), bil = read_b(kb, 3);
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][0] = mfma_bf16(ef[t].lo, brh, acc[t][0]);
                    acc[t][1] = mfma_bf16(ef[t].lo, bih, acc[t][1]);
                }
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][0] = mfma_bf16(ef[t].hi, brl, acc[t][0]);
                    acc[t][1] = mfma_bf16(ef[t].hi, bil, acc[t][1]);
                }
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    acc[t][0] = mfma_bf16(ef[t].hi, brh, acc[t][0]);
                    acc[t][1] = mfma_bf16(ef[t].hi, bih, acc[t][1]);
                }
            }
        }
    };
    // T <- split(X): lane writes rows 4q..4q+3 of its tiles in column lr
    auto store_t = [&](const f32x4 (&X)[RT][2]) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                bf16x8 hl;  // [0..3] hi, [4..7] lo
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const __bf16 h = (__bf16)X[t][c][r];
                    hl[r] = h;
                    hl[4 + r] = (__bf16)(X[t][c][r] - (float)h);
                }
                const uint4 u = __builtin_bit_cast(uint4, hl);
                const int k = (wave * RT + t) * 16 + 4 * q;
                *reinterpret_cast<uint2*>(Tt + ((2 * c + 0) * 16 + lr) * TP + k) = make_uint2(u.x, u.y);
                *reinterpret_cast<uint2*>(Tt + ((2 * c + 1) * 16 + lr) * TP + k) = make_uint2(u.z, u.w);
            }
    };
    // the walker's columns of item i (zero outside n x o), as accumulator rows
    auto load_cols = [&](double2 (&wx)[RT][4], int wv, int i) {
        const int ln = opaque(lane);
        const int wvo = opaque(wave);
        const int si = i == 0 ? 0 : s1;
        const int cl = (i == 0 ? 0 : c1) + (ln & 15);
        const bool ok = wv < nw;
        const double2* base = (si == 0 ? win_up : win_dn) + (ok ? (long long)wv * n * o : 0LL);
        const rsrc_t Wr = make_rsrc(base, ok ? (unsigned)n * o * 16u : 0u);
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = (wvo * RT + t) * 16 + 4 * (ln >> 4) + r;
                wx[t][r] = load_d2(Wr, cl < o ? (unsigned)(p * o + cl) * 16u : kOutside);
            }
    };
    auto to_acc = [&](f32x4 (&X)[RT][2], const double2 (&wx)[RT][4]) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                X[t][0][r] = (float)wx[t][r].x;
                X[t][1][r] = (float)wx[t][r].y;
            }
    };
    // output rows of item i of walker wv: written (rows < n, columns < o) and
    // read back (zero elsewhere: the walker's padding stays zero throughout),
    // through buffer offsets (out of range: a dropped store, a zero load), so
    // no store waits on a branch or a reloaded address
    auto out_rsrc = [&](int wv, int i) {
        const double2* p = ((i == 0 ? 0 : s1) == 0 ? wout_up : wout_dn) + (long long)wv * n * o;
        return make_rsrc(p, (unsigned)n * o * 16u);
    };
    auto out_off = [&](int i, int t, int r, int wv, int ln) {
        const int cl = (i == 0 ? 0 : c1) + (ln & 15);
        const int row = (wv * RT + t) * 16 + 4 * (ln >> 4) + r;
        return (cl < o && row < n) ? (unsigned)(row * o + cl) * 16u : kOutside;
    };
    auto park = [&](const rsrc_t& dst, int i, const f32x4 (&X)[RT][2], int wv, int ln) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double2 v = make_double2((double)X[t][0][r], (double)X[t][1][r]);
                __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, v), dst, out_off(i, t, r, wv, ln), 0, 0);
            }
    };
    auto unpark = [&](const rsrc_t& src, int i, f32x4 (&X)[RT][2], int wv, int ln) {
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double2 v = load_d2(src, out_off(i, t, r, wv, ln));
                X[t][0][r] = (float)v.x;
                X[t][1][r] = (float)v.y;
            }
    };
    // P += V . T for one k-block, A fragments (ar, ai) per row tile
    auto vt_block = [&](f32x4 (&P)[RT][2], int kb, const split8 (&ar)[RT], const split8 (&ai)[RT]) {
        const bf16x8 brh = read_b(kb, 0), brl = read_b(kb, 1);
        bf16x8 bih = read_b(kb, 2), bil = read_b(kb, 3);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][0] = mfma_bf16(ar[t].lo, brh, P[t][0]);
            P[t][1] = mfma_bf16(ar[t].lo, bih, P[t][1]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][0] = mfma_bf16(ar[t].hi, brl, P[t][0]);
            P[t][1] = mfma_bf16(ar[t].hi, bil, P[t][1]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][0] = mfma_bf16(ar[t].hi, brh, P[t][0]);
            P[t][1] = mfma_bf16(ar[t].hi, bih, P[t][1]);
        }
        bih = neg8(bih);
        bil = neg8(bil);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][1] = mfma_bf16(ai[t].lo, brh, P[t][1]);
            P[t][0] = mfma_bf16(ai[t].lo, bih, P[t][0]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][1] = mfma_bf16(ai[t].hi, brl, P[t][1]);
            P[t][0] = mfma_bf16(ai[t].hi, bil, P[t][0]);
        }
#pragma unroll
        for (int t = 0; t < RT; ++t) {
            P[t][1] = mfma_bf16(ai[t].hi, brh, P[t][1]);
            P[t][0] = mfma_bf16(ai[t].hi, bih, P[t][0]);
        }
    };

    float er[RT][KMAX][8];
    double2 wa[RT][4], wb[RT][4];
    int w = blockIdx.x;
    load_cols(wa, w, 0);
    load_cols(wb, w, 1);
    if constexpr (MODE != 2) load_e(Er, er, row0, kq);
#pragma unroll 1
    for (; w < nw; w += gridDim.x) {
        const int wn = w + gridDim.x;
        const rsrc_t out0 = out_rsrc(w, 0);
        const rsrc_t out1 = out_rsrc(w, 1);
        if constexpr (MODE == 2) {
            const int ln = opaque(lane);
            const int wvo = opaque(wave);
            load_e(Er, er, wvo * RT * 16 + (ln & 15), 8 * (ln >> 4));
        }
        f32x4 out[RT][2];
        // ---- E W of both items ----
        {
            f32x4 acc[RT][2];
            to_acc(acc, wa);
            __syncthreads();  // the previous walker's reads of T are done
            store_t(acc);
            __syncthreads();
            e_product(acc, er, row0, kq);
            f32x4 acc1[RT][2];
            to_acc(acc1, wb);
            __syncthreads();
            store_t(acc1);
            __syncthreads();
            e_product(acc1, er, row0, kq);
            park(out1, 1, acc1, wave, lane);
#pragma unroll
            for (int t = 0; t < RT; ++t) {
                out[t][0] = acc[t][0];
                out[t][1] = acc[t][1];
            }
            __syncthreads();
            store_t(acc);
        }
        // ---- V resident: the LDS k-blocks first, then the register k-blocks
        // (lane-derived offsets recomputed per walker from an opaque lane id,
        // so none is kept, and spilled, across the walker loop) ----
        const int lnv = opaque(lane);
        const int wvv = opaque(wave);
        const int lrv = lnv & 15, qv = lnv >> 4, kqv = 8 * qv, row0v = wvv * RT * 16 + lrv;
        const rsrc_t Vr = make_rsrc(vhs + (long long)w * vrow, (unsigned)vrow * 4u);
        const rsrc_t Vi = make_rsrc(vhs + vplane + (long long)w * vrow, (unsigned)vrow * 4u);
        if constexpr (NL > 0) {
            float xl[RT][NL][2][8];
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    const int k0 = (KR + l) * 32 + kqv;
                    if constexpr (SYMV) {
                        load_frag_sym(Vr, nbv, wvv * RT + t, KR + l, lrv, qv, xl[t][l][0]);
                        load_frag_sym(Vi, nbv, wvv * RT + t, KR + l, lrv, qv, xl[t][l][1]);
                    } else {
                        load_row8(Vr, ldv, n, row0v + 16 * t, k0, xl[t][l][0]);
                        load_row8(Vi, ldv, n, row0v + 16 * t, k0, xl[t][l][1]);
                    }
                }
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    uint4* dst = VL + ((size_t)((wvv * RT + t) * NL + l) * 4) * 64 + lnv;
                    const split8 a = SYMV ? unpack_pairs(xl[t][l][0]) : split_f32x8(xl[t][l][0]);
                    const split8 c = SYMV ? unpack_pairs(xl[t][l][1]) : split_f32x8(xl[t][l][1]);
                    dst[0 * 64] = __builtin_bit_cast(uint4, a.hi);
                    dst[1 * 64] = __builtin_bit_cast(uint4, a.lo);
                    dst[2 * 64] = __builtin_bit_cast(uint4, c.hi);
                    dst[3 * 64] = __builtin_bit_cast(uint4, c.lo);
                }
        }
        split8 vr[RT][KR], vi[RT][KR];
        {
            float xv[RT][KR][2][8];
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int kb = 0; kb < KR; ++kb) {
                    const int k0 = kb * 32 + kqv;
                    if constexpr (SYMV) {
                        load_frag_sym(Vr, nbv, wvv * RT + t, kb, lrv, qv, xv[t][kb][0]);
                        load_frag_sym(Vi, nbv, wvv * RT + t, kb, lrv, qv, xv[t][kb][1]);
                    } else {
                        load_row8(Vr, ldv, n, row0v + 16 * t, k0, xv[t][kb][0]);
                        load_row8(Vi, ldv, n, row0v + 16 * t, k0, xv[t][kb][1]);
                    }
                }
#pragma unroll
            for (int t = 0; t < RT; ++t)
#pragma unroll
                for (int kb = 0; kb < KR; ++kb) {
                    if constexpr (SYMV) {
                        vr[t][kb] = unpack_pairs(xv[t][kb][0]);
                        vi[t][kb] = unpack_pairs(xv[t][kb][1]);
                    } else {
                        vr[t][kb] = split_f32x8(xv[t][kb][0]);
                        vi[t][kb] = split_f32x8(xv[t][kb][1]);
                    }
                }
        }
        __syncthreads();

        // ---- Taylor terms: T_k = (i sqrt(dt) / k) V T_{k-1}, out += T_k ----
        auto taylor = [&]() {
#pragma unroll 1
            for (int k = 1; k <= n_terms; ++k) {
                f32x4 P[RT][2];
#pragma unroll
                for (int t = 0; t < RT; ++t) P[t][0] = P[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < KR; ++kb) {
                    if (kb < nk) {
                        split8 ar[RT], ai[RT];
#pragma unroll
                        for (int t = 0; t < RT; ++t) {
                            ar[t] = vr[t][kb];
                            ai[t] = vi[t][kb];
                        }
                        vt_block(P, kb, ar, ai);
                    }
                }
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    split8 ar[RT], ai[RT];
#pragma unroll
                    for (int t = 0; t < RT; ++t) {
                        const uint4* src = VL + ((size_t)((wave * RT + t) * NL + l) * 4) * 64 + lane;
                        ar[t].hi = __builtin_bit_cast(bf16x8, src[0 * 64]);
                        ar[t].lo = __builtin_bit_cast(bf16x8, src[1 * 64]);
                        ai[t].hi = __builtin_bit_cast(bf16x8, src[2 * 64]);
                        ai[t].lo = __builtin_bit_cast(bf16x8, src[3 * 64]);
                    }
                    vt_block(P, KR + l, ar, ai);
                }
                const float ck = (float)(sqrt_dt / (double)k);
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    const f32x4 re = -ck * P[t][1];
                    const f32x4 im = ck * P[t][0];
                    P[t][0] = re;
                    P[t][1] = im;
                    out[t][0] += re;
                    out[t][1] += im;
                }
                __syncthreads();
                store_t(P);
                __syncthreads();
            }
        };
        taylor();
        {
            // first sum parked; the second item's E W back as its sum's start
            const int lane1 = opaque(lane);
            const int wave1 = opaque(wave);
            park(out0, 0, out, wave1, lane1);
            unpark(out1, 1, out, wave1, lane1);
            __syncthreads();
            store_t(out);
            __syncthreads();
        }
        taylor();

        // ---- W_out = E out = out + (E - I) out for both items; the next
        // walker's columns in flight (indices recomputed from an opaque lane id:
        // nothing address-like stays live across the Taylor loops) ----
        const int lane2 = opaque(lane);
        const int wave2 = opaque(wave);
        const int r02 = wave2 * RT * 16 + (lane2 & 15);
        const int k82 = 8 * (lane2 >> 4);
        if constexpr (MODE != 0) {
            load_e(ErL, er, r02, k82);
            load_cols(wa, wn, 0);
            load_cols(wb, wn, 1);
            __syncthreads();
            store_t(out);
            __syncthreads();
            e_product(out, er, r02, k82);
            {
                // (the load before the store: a load issued after a store waits
                // for it, the counters being in order)
                f32x4 nx[RT][2];
                unpark(out0, 0, nx, wave2, lane2);
                park(out1, 1, out, wave2, lane2);
#pragma unroll
                for (int t = 0; t < RT; ++t) {
                    out[t][0] = nx[t][0];
                    out[t][1] = nx[t][1];
                }
            }
            __syncthreads();
            store_t(out);
            __syncthreads();
            e_product(out, er, r02, k82);
            park(out0, 0, out, wave2, lane2);
        } else {
            // the first item's sum already sits in its output rows (parked
            // after the first series); the second's goes to its rows, after
            // the next walker's loads are issued (so they do not wait for it)
            load_e(Er, er, r02, k82);
            load_cols(wa, wn, 0);
            load_cols(wb, wn, 1);
            park(out1, 1, out, wave2, lane2);
        }
    }
}

size_t bf16x3_lds(int n, int nl) {
    const int NP = afq::round_up(n, 16);
    const int nk = afq::ceil_div(NP, 32);
    const int nwave = afq::ceil_div(NP / 16, RT);
    return (size_t)4 * 16 * (nk * 32 + 8) * 2 + (size_t)nwave * RT * nl * 4 * 64 * 16;
}

// E: the first half's matrix; Elast: the final half's (nullptr: no final half)
template <int NL, bool SYMV>
int launch_bf16x3(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                  const float* E, const float* Elast, const float* vhs, const double2* wu, const double2* wd,
                  double2* ou, double2* od, hipStream_t st) {
    const int NP = afq::round_up(n, 16);
    const int nct = afq::ceil_div(o, 16);
    const int nwave = afq::ceil_div(NP / 16, RT);
    const size_t lds = bf16x3_lds(n, NL);
    const int group_map = (nw % 8 == 0) ? 1 : 0;
    const int nb = NP / 16;
    const long long vplane = SYMV ? (long long)nw * nb * (nb + 1) / 2 * 256 : (long long)nw * n * ldv;
    // upper-block V, two items per walker and a walker for every CU: the
    // persistent pair kernel
    bool pair = false;
    if constexpr (SYMV) {
        if (nspin * nct == 2 && nw >= afq::cu_count()) {
            pair = true;
            const dim3 grid((unsigned)afq::cu_count()), block(64 * nwave);
#define AFQ_PAIR(M)                                                                                             \
    do {                                                                                                        \
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_bf16x3_pair_kernel<NL, M>);               \
        hipLaunchKernelGGL((trotter_bf16x3_pair_kernel<NL, M>), grid, block, lds, st, nw, n, o, nspin, nct,     \
                           n_terms, sqrt_dt, E, Elast, vhs, vplane, wu, wd, ou, od);                            \
    } while (0)
            if (!Elast)
                AFQ_PAIR(0);
            else if (Elast == E)
                AFQ_PAIR(1);
            else
                AFQ_PAIR(2);
#undef AFQ_PAIR
        }
    }
    if (!pair) {
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_bf16x3_kernel<NL, SYMV>);
        hipLaunchKernelGGL((trotter_bf16x3_kernel<NL, SYMV>), dim3((unsigned)(nw * nspin * nct)), dim3(64 * nwave),
                           lds, st, nw, n, o, nspin, nct, ldv, n_terms, sqrt_dt, E, vhs, vplane, wu, wd, ou, od,
                           group_map, 0, nullptr, Elast, Elast ? 1 : 0);
    }
    return afq::check_launch(SYMV ? "afq_trotter_bf16x3_sym" : "afq_trotter_bf16x3");
}

template <bool SYMV>
int launch_bf16x3_nl(int nl, int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                     const float* E, const float* Elast, const float* vhs, const double* win_up,
                     const double* win_dn, double* wout_up, double* wout_dn, hipStream_t st) {
    const double2* wu = (const double2*)win_up;
    const double2* wd = (const double2*)win_dn;
    double2* ou = (double2*)wout_up;
    double2* od = (double2*)wout_dn;
    if (nl == 0)
        return launch_bf16x3<0, SYMV>(nw, n, o, nspin, ldv, n_terms, sqrt_dt, E, Elast, vhs, wu, wd, ou, od, st);
    if (nl == 1)
        return launch_bf16x3<1, SYMV>(nw, n, o, nspin, ldv, n_terms, sqrt_dt, E, Elast, vhs, wu, wd, ou, od, st);
    return launch_bf16x3<2, SYMV>(nw, n, o, nspin, ldv, n_terms, sqrt_dt, E, Elast, vhs, wu, wd, ou, od, st);
}

// the checks both entry points share; the k-blocks of V kept in LDS in *nl
int check_bf16x3(const char* name, int nw, int n, int o, int nspin, int ldv, int n_terms,
                 const float* exp_h1_half, const float* vhs, const double* win_up, const double* win_dn,
                 double* wout_up, double* wout_dn, int* nl) {
    AFQ_REQUIRE(nw > 0 && n > 0 && o > 0 && o <= n, "%s: bad sizes nw=%d n=%d o=%d", name, nw, n, o);
    AFQ_REQUIRE(nspin == 1 || nspin == 2, "%s: nspin must be 1 or 2", name);
    AFQ_REQUIRE(ldv >= n && ldv % 4 == 0, "%s: ldv=%d must be >= n and a multiple of 4", name, ldv);
    AFQ_REQUIRE(n_terms >= 0, "%s: n_terms < 0", name);
    AFQ_REQUIRE(exp_h1_half && vhs && win_up && wout_up, "%s: null pointer", name);
    AFQ_REQUIRE(nspin == 1 || (win_dn && wout_dn), "%s: null spin-down pointer", name);
    AFQ_REQUIRE(win_up != wout_up && (nspin == 1 || win_dn != wout_dn), "%s: output aliases input", name);
    AFQ_REQUIRE((long long)nw * nspin * afq::ceil_div(o, 16) < (1LL << 31), "%s: too many walkers", name);
    const int NP = afq::round_up(n, 16);
    const int nk = afq::ceil_div(NP, 32);
    *nl = nk > KR ? nk - KR : 0;
    return 0;
}

}  // namespace

// Declared in afqmc_hip.h.  Sizes beyond the on-chip capacity of the split
// kernel (n > 224) go to the fp32 step, which is more accurate.
extern "C" int afq_trotter_bf16x3(int nw, int n, int o, int nspin, int ldv, int n_terms,
                                  double sqrt_dt, const float* exp_h1_half, const float* vhs,
                                  const double* win_up, const double* win_dn, double* wout_up,
                                  double* wout_dn, void* stream) {
    int nl = 0;
    const int rc = check_bf16x3("afq_trotter_bf16x3", nw, n, o, nspin, ldv, n_terms, exp_h1_half, vhs, win_up,
                                win_dn, wout_up, wout_dn, &nl);
    if (rc != 0) return rc;
    if (nl > NL_MAX || bf16x3_lds(n, nl) > 160 * 1024)
        return afq_trotter_f32(nw, n, o, nspin, ldv, n_terms, sqrt_dt, exp_h1_half, vhs, win_up, win_dn,
                               wout_up, wout_dn, stream);
    return launch_bf16x3_nl<false>(nl, nw, n, o, nspin, ldv, n_terms, sqrt_dt, exp_h1_half, exp_h1_half, vhs,
                                   win_up, win_dn, wout_up, wout_dn, (hipStream_t)stream);
}

namespace {

int trotter_bf16x3_sym(const char* name, int nw, int n, int o, int nspin, int n_terms, double sqrt_dt,
                       const unsigned* e_first, const unsigned* e_last, const unsigned* vhs_sym,
                       const double* win_up, const double* win_dn, double* wout_up, double* wout_dn,
                       void* stream) {
    int nl = 0;
    const int ldv = afq::round_up(n, 16);  // unused by this layout
    const int rc = check_bf16x3(name, nw, n, o, nspin, ldv, n_terms, (const float*)e_first,
                                (const float*)vhs_sym, win_up, win_dn, wout_up, wout_dn, &nl);
    if (rc != 0) return rc;
    AFQ_REQUIRE(nl <= NL_MAX && bf16x3_lds(n, nl) <= 160 * 1024, "%s: n=%d exceeds the on-chip maximum of 224",
                name, n);
    return launch_bf16x3_nl<true>(nl, nw, n, o, nspin, ldv, n_terms, sqrt_dt, (const float*)e_first,
                                  (const float*)e_last, (const float*)vhs_sym, win_up, win_dn, wout_up, wout_dn,
                                  (hipStream_t)stream);
}

}  // namespace

// Declared in afqmc_hip.h: V and E as upper 16 x 16 blocks; n <= 224 only
// (there is no fp32 kernel for this layout).
extern "C" int afq_trotter_bf16x3_sym(int nw, int n, int o, int nspin, int n_terms, double sqrt_dt,
                                      const unsigned* exp_h1_half_sym, const unsigned* vhs_sym,
                                      const double* win_up, const double* win_dn, double* wout_up,
                                      double* wout_dn, void* stream) {
    return trotter_bf16x3_sym("afq_trotter_bf16x3_sym", nw, n, o, nspin, n_terms, sqrt_dt, exp_h1_half_sym,
                              exp_h1_half_sym, vhs_sym, win_up, win_dn, wout_up, wout_dn, stream);
}

// Declared in afqmc_hip.h: the same step with its first half on e_first and
// its last on e_last (NULL: no last half).
extern "C" int afq_trotter_bf16x3_sym_halves(int nw, int n, int o, int nspin, int n_terms, double sqrt_dt,
                                             const unsigned* e_first, const unsigned* e_last,
                                             const unsigned* vhs_sym, const double* win_up, const double* win_dn,
                                             double* wout_up, double* wout_dn, void* stream) {
    return trotter_bf16x3_sym("afq_trotter_bf16x3_sym_halves", nw, n, o, nspin, n_terms, sqrt_dt, e_first, e_last,
                              vhs_sym, win_up, win_dn, wout_up, wout_dn, stream);
}

namespace {

template <bool SYMV>
int trotter_bf16x3_grouped(const char* name, int count, const afq_system_step* sys, int total_blocks, int max_n,
                           int n_terms, hipStream_t st, int first_full = 0, int last_half = 1) {
    AFQ_REQUIRE(count > 0 && total_blocks > 0, "%s: count=%d total_blocks=%d", name, count, total_blocks);
    AFQ_REQUIRE(sys, "%s: null pointer", name);
    AFQ_REQUIRE(max_n > 0 && max_n <= 16 * 2 * (KR + NL_MAX), "%s: max_n=%d not in 1..%d", name, max_n,
                16 * 2 * (KR + NL_MAX));
    AFQ_REQUIRE(n_terms >= 0, "%s: n_terms < 0", name);
    const int NP = afq::round_up(max_n, 16);
    const int nk = afq::ceil_div(NP, 32);
    const int nl = nk > KR ? nk - KR : 0;
    const size_t lds = bf16x3_lds(max_n, nl);
    AFQ_REQUIRE(lds <= 160 * 1024, "%s: max_n=%d needs %zu bytes of LDS", name, max_n, lds);
    const dim3 block(64 * afq::ceil_div(NP / 16, RT));
    const dim3 grid((unsigned)total_blocks);
#define AFQ_TROTTER_GROUPED(NLV)                                                                                  \
    do {                                                                                                          \
        if (lds > 64 * 1024) afq::allow_full_lds((const void*)trotter_bf16x3_kernel<NLV, SYMV, true>);            \
        hipLaunchKernelGGL((trotter_bf16x3_kernel<NLV, SYMV, true>), grid, block, lds, st, 0, 0, 0, 1, 0, 0,         \
                           n_terms, 0.0, nullptr, nullptr, 0LL, nullptr, nullptr, nullptr, nullptr, 0, count, sys,  \
                           nullptr, last_half, first_full);                                                         \
    } while (0)
    if (nl == 0)
        AFQ_TROTTER_GROUPED(0);
    else if (nl == 1)
        AFQ_TROTTER_GROUPED(1);
    else
        AFQ_TROTTER_GROUPED(2);
#undef AFQ_TROTTER_GROUPED
    return afq::check_launch(name);
}

}  // namespace

extern "C" int afq_trotter_bf16x3_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n,
                                          int n_terms, void* stream) {
    return trotter_bf16x3_grouped<false>("afq_trotter_bf16x3_grouped", count, sys, total_blocks, max_n, n_terms,
                                         (hipStream_t)stream);
}

extern "C" int afq_trotter_bf16x3_sym_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n,
                                              int n_terms, void* stream) {
    return trotter_bf16x3_grouped<true>("afq_trotter_bf16x3_sym_grouped", count, sys, total_blocks, max_n, n_terms,
                                        (hipStream_t)stream);
}

extern "C" int afq_trotter_bf16x3_sym_grouped_halves(int count, const afq_system_step* sys, int total_blocks,
                                                     int max_n, int n_terms, int first_full, int last_half,
                                                     void* stream) {
    return trotter_bf16x3_grouped<true>("afq_trotter_bf16x3_sym_grouped_halves", count, sys, total_blocks, max_n,
                                        n_terms, (hipStream_t)stream, first_full, last_half);
}
