// Split-bf16 ("bf16x3") GEMM and the operand converters of the mixed step.
//
// An fp64 operand x is carried as two bf16 planes, hi = bf16(x) and
// lo = bf16(x - hi), so hi + lo holds ~16 significant bits.  The product
// A B^T is accumulated in fp32 as  Ahi Bhi + Ahi Blo + Alo Bhi  (the dropped
// Alo Blo term is ~2^-16 relative), three v_mfma_f32_16x16x32_bf16 per
// 16x16x32 step.
//
// Tiling: 128 x 128 x 32 per 256-thread workgroup (2 x 2 waves of 64 x 64),
// all four operand planes staged through double-buffered LDS (80 KB, two
// workgroups per CU).  Each lane reads 8 consecutive k of one row (16 bytes)
// for the A and B fragments, so A and B see the same k order.  K must be a
// multiple of 8 (operand planes are zero-padded to that by the converter).
#include <cstdint>

#include "afqmc_hip.h"
#include "common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 32;
constexpr int PITCH = 40;  // 80-byte rows: conflict-free 16-byte LDS reads
constexpr int THREADS = 256;
constexpr int GROUP_M = 8;


// SYM: the columns are symmetric-VHS block entries and the tile is stored
// expanded (afq::sym_store); C is then float, ldc unused.
// Split K: blockIdx.y = kz takes k in [kz kc, (kz + 1) kc) (kc a multiple of
// 32) and writes its partial product to C + kz c_stride.
// One 128 x 128 tile of C (rows m0.., columns n0..) over the whole K, staged
// through the workgroup's double-buffered LDS planes S.
template <typename TOut, bool SYM>
__device__ __forceinline__ void bf16x3_tile(int M, int N, int K, float alpha, const uint16_t* __restrict__ Ahi,
                                            const uint16_t* __restrict__ Alo, long long lda,
                                            const uint16_t* __restrict__ Bhi, const uint16_t* __restrict__ Blo,
                                            long long ldb, TOut* __restrict__ C, long long ldc, int m0, int n0,
                                            afq::sym_dst sd, uint16_t (&S)[2][4][BM * PITCH]) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1;
    const int wn = wave & 1;

    uint4 reg[4][2];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const uint16_t* src = a == 0 ? Ahi : a == 1 ? Alo : a == 2 ? Bhi : Blo;
            const int base = a < 2 ? m0 : n0;
            const int lim = a < 2 ? M : N;
            const long long ld = a < 2 ? lda : ldb;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int chunk = tid + THREADS * j;
                const int row = chunk >> 2;
                const int kc = (chunk & 3) * 8;
                const int grow = base + row;
                const int gk = k0 + kc;
                if (grow < lim && gk < K)
                    reg[a][j] = *reinterpret_cast<const uint4*>(src + (long long)grow * ld + gk);
                else
                    reg[a][j] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int chunk = tid + THREADS * j;
                const int row = chunk >> 2;
                const int kc = (chunk & 3) * 8;
                *reinterpret_cast<uint4*>(&S[buf][a][row * PITCH + kc]) = reg[a][j];
            }
    };

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nk = afq::ceil_div(K, BK);
    load_tile(0);
    store_tile(0);
    __syncthreads();

    const int lr = lane & 15;
    const int lk = (lane >> 4) * 8;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tile((kt + 1) * BK);
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int off = (wm * 64 + i * 16 + lr) * PITCH + lk;
            ah[i] = *reinterpret_cast<const bf16x8*>(&S[buf][0][off]);
            al[i] = *reinterpret_cast<const bf16x8*>(&S[buf][1][off]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int off = (wn * 64 + j * 16 + lr) * PITCH + lk;
            bh[j] = *reinterpret_cast<const bf16x8*>(&S[buf][2][off]);
            bl[j] = *reinterpret_cast<const bf16x8*>(&S[buf][3][off]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
            }
        if (kt + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }

    if constexpr (SYM) {
        // a 128-column tile lies inside one block; the wave's 4 column tiles
        // are 4 consecutive block rows
        int bi, bj;
        afq::sym_block(n0 >> 8, sd.nb, bi, bj);
        const int pi0 = ((n0 & 255) >> 4) + wn * 4;
        const long long ldf = (long long)sd.n * sd.ldv;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
                if (row >= M) continue;
                float v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) v[j] = alpha * acc[i][j][r];
                afq::sym_store<4>((float*)C + row * ldf, sd, bi, bj, pi0, lr, v);
            }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
                if (row < M && col < N) C[(long long)row * ldc + col] = (TOut)(alpha * acc[i][j][r]);
            }
        }
}

// lower_b (whole K only): row j of B is zero past column j, so the tile of
// columns n0.. stops its k-loop at k = n0 + BN (the zeros inside that range
// are still read); the column tiles are then taken in afq::zigzag order, so
// that every XCD's contiguous share of tiles holds the same work
template <typename TOut, bool SYM = false>
__global__ __launch_bounds__(THREADS, 2) void gemm_bf16x3_nt_kernel(
    int M, int N, int K, float alpha, const uint16_t* __restrict__ Ahi,
    const uint16_t* __restrict__ Alo, long long lda, const uint16_t* __restrict__ Bhi,
    const uint16_t* __restrict__ Blo, long long ldb, TOut* __restrict__ C, long long ldc,
    int tiles_m, int tiles_n, afq::sym_dst sd = afq::sym_dst{0, 0, 0}, int kc = 0, long long c_stride = 0,
    int lower_b = 0) {
    if (gridDim.y > 1) {
        const int k_begin = blockIdx.y * kc;
        K = min(K - k_begin, kc);
        Ahi += k_begin;
        Alo += k_begin;
        Bhi += k_begin;
        Blo += k_begin;
        C += blockIdx.y * c_stride;
    }
    __shared__ __attribute__((aligned(16))) uint16_t S[2][4][BM * PITCH];

    const int total = tiles_m * tiles_n;
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int per_group = GROUP_M * tiles_n;
    const int group = t / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int in_group = t - group * per_group;
    const int m0 = (first_m + in_group % gsize) * BM;
    const int n0 = (lower_b ? afq::zigzag(in_group / gsize, tiles_n) : in_group / gsize) * BN;
    if (lower_b) K = min(K, n0 + BN);

    bf16x3_tile<TOut, SYM>(M, N, K, alpha, Ahi, Alo, lda, Bhi, Blo, ldb, C, ldc, m0, n0, sd, S);
}

// Grouped bf16x3 GEMM: workgroup -> (problem, 128 x 128 tile) by the
// problems' tile_start prefix sums, the linear workgroup id first made
// XCD-contiguous (each XCD runs a contiguous range of tiles and shares their
// panels in its L2).
__global__ __launch_bounds__(THREADS, 2) void gemm_bf16x3_grouped_kernel(int count, int total,
                                                                         const afq_gemm_problem_bf16x3* __restrict__ P) {
    __shared__ __attribute__((aligned(16))) uint16_t S[2][4][BM * PITCH];
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int pi = afq::find_group(count, t, [&](int i) { return P[i].tile_start; });
    const afq_gemm_problem_bf16x3& p = P[pi];
    const int tiles_m = afq::ceil_div(p.M, BM);
    const int local = t - p.tile_start;
    const int m0 = (local % tiles_m) * BM;
    const int nn = local / tiles_m;
    // b_lower: column tiles in zigzag order, each k-loop stopping at its last column
    const int n0 = (p.b_lower ? afq::zigzag(nn, afq::ceil_div(p.N, BN)) : nn) * BN;
    const int K = p.b_lower ? min(p.K, n0 + BN) : p.K;
    if (p.sym_n > 0)
        bf16x3_tile<float, true>(p.M, p.N, K, p.alpha, p.A_hi, p.A_lo, p.lda, p.B_hi, p.B_lo, p.ldb, (float*)p.C,
                                 0LL, m0, n0, afq::sym_dst{p.sym_n, p.sym_ldv, afq::ceil_div(p.sym_n, 16)}, S);
    else if (p.c_fp64 == 2)
        bf16x3_tile<afq::bf16_pair, false>(p.M, p.N, K, p.alpha, p.A_hi, p.A_lo, p.lda, p.B_hi, p.B_lo, p.ldb,
                                           (afq::bf16_pair*)p.C, p.ldc, m0, n0, afq::sym_dst{0, 0, 0}, S);
    else if (p.c_fp64)
        bf16x3_tile<double, false>(p.M, p.N, K, p.alpha, p.A_hi, p.A_lo, p.lda, p.B_hi, p.B_lo, p.ldb,
                                   (double*)p.C, p.ldc, m0, n0, afq::sym_dst{0, 0, 0}, S);
    else
        bf16x3_tile<float, false>(p.M, p.N, K, p.alpha, p.A_hi, p.A_lo, p.lda, p.B_hi, p.B_lo, p.ldb,
                                  (float*)p.C, p.ldc, m0, n0, afq::sym_dst{0, 0, 0}, S);
}

// Large-problem variant: 256 x 256 x 32 per 512-thread workgroup (4 x 2
// waves of 64 x 128), so each LDS fragment feeds twice the MFMAs of the
// 128 x 128 tile.  The four operand planes of a k-step (64 KB) go from global
// memory straight to LDS (buffer_load ... lds, 16 bytes per lane, no staging
// registers), double-buffered (128 KB, one workgroup per CU).  Rows are 64
// bytes, unpadded; the 16-byte chunk c of row r is stored at position
// c ^ ((r >> 1) & 3), which makes the fragment reads conflict-free.  Rows past
// M / N and k-chunks past K fall outside the buffer descriptors and read as
// zero.
constexpr int BM2 = 256;
constexpr int BN2 = 256;
constexpr int THREADS2 = 512;
constexpr int GROUP_M2 = 8;
constexpr unsigned kOutside = 0x80000000u;

typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef __attribute__((address_space(3))) void* lds_ptr_t;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, unsigned long long bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0,
                                            (int)(bytes > 0x7fffffffull ? 0x7fffffffull : bytes), 0x00020000);
}

// Tile geometry of the LDS-DMA main loop: BM = 64 WM rows x 256 columns of C
// per workgroup of 2 WM waves (WM x 2 waves of 64 x 128 each).  A k-step
// moves the four operand planes (A hi, A lo: BM rows; B hi, B lo: 256 rows;
// 64 bytes per row) in 1 KB blocks of 16 rows, BPW blocks per wave.
// With WM = 4 each wave's blocks lie in one plane (waves 0..3 move A, 4..7 B)
// and the B planes run three k-steps deep (2 x 32 + 3 x 32 KB = the CU's
// 160 KB of LDS): the B movers issue two k-steps ahead, the A movers one.
template <int WM>
struct dma_shape {
    static constexpr int BM = 64 * WM;
    static constexpr int THREADS = 128 * WM;
    static constexpr int ABLK = BM / 16;            // blocks per A plane
    static constexpr int NBLK = 2 * ABLK + 2 * 16;  // blocks per k-step
    static constexpr int BPW = NBLK / (2 * WM);     // blocks per wave
    static constexpr int NB = WM == 4 ? 3 : 2;      // B stages
    static constexpr int A_STAGE = 2 * ABLK * 64;   // uint4 of one k-step's A planes
    static constexpr int B_STAGE = 2 * 16 * 64;     // uint4 of one k-step's B planes
    static constexpr int SIZE = 2 * A_STAGE + NB * B_STAGE;
    // first uint4 of plane p (A hi, A lo, B hi, B lo) of k-step kt
    __host__ __device__ static constexpr int plane_at(int p, int kt) {
        return p < 2 ? (kt & 1) * A_STAGE + p * ABLK * 64
                     : 2 * A_STAGE + (kt % NB) * B_STAGE + (p - 2) * 16 * 64;
    }
};

// acc[i][j] (rows wm*64 + 16 i, columns wn*128 + 16 j of the tile at m0, n0)
// = A B^T over the whole K, with S the staged planes (SIZE uint4).
template <int WM>
__device__ __forceinline__ void dma_mainloop(int M, int N, int K, const uint16_t* __restrict__ Ahi,
                                             const uint16_t* __restrict__ Alo, long long lda,
                                             const uint16_t* __restrict__ Bhi, const uint16_t* __restrict__ Blo,
                                             long long ldb, int m0, int n0, uint4* S, f32x4 (&acc)[4][8]) {
    using D = dma_shape<WM>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // wave-uniform (scalar): the descriptors and LDS bases below must be
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave % WM;  // 64-row slice of the tile
    const int wn = wave / WM;  // 128-column slice

    // copy assignment: block gb = wave * BPW + i of the k-step, in plane order
    // (A hi, A lo, B hi, B lo); lane L fills slot L of a block: row L / 4,
    // position L % 4, i.e. chunk (L % 4) ^ ((L / 8) & 3)
    const int r_loc = lane >> 2;
    const int chunk = (lane & 3) ^ ((lane >> 3) & 3);
    // the descriptor of plane p (A planes from row m0, B planes from row n0)
    auto plane_rsrc = [&](int p) {
        const uint16_t* src = p == 0 ? Ahi + (long long)m0 * lda
                            : p == 1 ? Alo + (long long)m0 * lda
                            : p == 2 ? Bhi + (long long)n0 * ldb
                                     : Blo + (long long)n0 * ldb;
        const long long ld = p < 2 ? lda : ldb;
        const int rows_left = max(p < 2 ? M - m0 : N - n0, 0);
        return make_rsrc(src, (unsigned long long)rows_left * ld * 2);
    };
    // WM = 4: each wave's eight blocks lie in one plane (one descriptor per wave)
    const int plane4 = wave >> 1;
    const rsrc_t R4 = plane_rsrc(WM == 4 ? plane4 : 0);
    const rsrc_t R0 = WM == 4 ? R4 : plane_rsrc(0);
    const rsrc_t R1 = WM == 4 ? R4 : plane_rsrc(1);
    const rsrc_t R2 = WM == 4 ? R4 : plane_rsrc(2);
    const rsrc_t R3 = WM == 4 ? R4 : plane_rsrc(3);
    auto issue_stage = [&](int kt) {
        const int k = kt * 32 + chunk * 8;
        // k >= K sets the top bit: the offset lands outside the descriptor
        const unsigned past = (unsigned)((K - 1 - k) >> 31) & kOutside;
#pragma unroll
        for (int i = 0; i < D::BPW; ++i) {
            int plane, blk;
            rsrc_t R;
            if constexpr (WM == 4) {
                plane = plane4;
                blk = (wave & 1) * 8 + i;
                R = R4;
            } else {
                const int gb = wave * D::BPW + i;
                plane = gb < D::ABLK ? 0 : gb < 2 * D::ABLK ? 1 : gb < 2 * D::ABLK + 16 ? 2 : 3;
                blk = gb - (plane < 2 ? plane * D::ABLK : 2 * D::ABLK + (plane - 2) * 16);
                R = plane == 0 ? R0 : plane == 1 ? R1 : plane == 2 ? R2 : R3;
            }
            const long long ld = plane < 2 ? lda : ldb;
            const int row = blk * 16 + r_loc;
            const unsigned off = (unsigned)(((long long)row * ld + k) * 2) | past;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(R, (lds_ptr_t)&S[D::plane_at(plane, kt) + blk * 64], 16, off,
                                                     0, 0, 0);
        }
    };

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // fragment read: row lr of a 16-row tile, chunk q at its swizzled position
    const int lr = lane & 15;
    const int q = lane >> 4;
    const int lane_slot = lr * 4 + (q ^ ((lr >> 1) & 3));

    const int nk = afq::ceil_div(K, 32);
    // a wave that moves B with three B stages runs two k-steps ahead: it keeps
    // its eight loads of the later k-step in flight at each wait
    const bool deep = D::NB == 3 && plane4 >= 2;
    issue_stage(0);
    if (deep && nk > 1) {
        issue_stage(1);
        asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
    } else {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int ka = deep ? kt + 2 : kt + 1;  // the k-step this wave issues now
        const bool ahead = ka < nk;
        if (ahead) issue_stage(ka);
        const uint4* A0 = S + D::plane_at(0, kt);
        const uint4* A1 = S + D::plane_at(1, kt);
        const uint4* B0 = S + D::plane_at(2, kt);
        const uint4* B1 = S + D::plane_at(3, kt);
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int slot = (wm * 64 + i * 16) * 4 + lane_slot;
            ah[i] = __builtin_bit_cast(bf16x8, A0[slot]);
            al[i] = __builtin_bit_cast(bf16x8, A1[slot]);
        }
#pragma unroll
        for (int jh = 0; jh < 2; ++jh) {
            bf16x8 bh[4], bl[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int slot = (wn * 128 + (jh * 4 + j) * 16) * 4 + lane_slot;
                bh[j] = __builtin_bit_cast(bf16x8, B0[slot]);
                bl[j] = __builtin_bit_cast(bf16x8, B1[slot]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    f32x4& c = acc[i][jh * 4 + j];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], c, 0, 0, 0);
                }
        }
        // k-step kt + 1 has landed in LDS and every wave is done with this one
        if (deep && ahead)
            asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }
}

// Plain store of a DMA tile: element (row, col) = alpha acc, inside M x N.
template <int WM, typename TOut>
__device__ __forceinline__ void dma_store(int M, int N, float alpha, TOut* __restrict__ C, long long ldc, int m0,
                                          int n0, const f32x4 (&acc)[4][8]) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave % WM, wn = wave / WM;
    const int lr = lane & 15, q = lane >> 4;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int col = n0 + wn * 128 + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + q * 4 + r;
                if (row < M && col < N) C[(long long)row * ldc + col] = (TOut)(alpha * acc[i][j][r]);
            }
        }
}

// Split K and lower_b as in gemm_bf16x3_nt_kernel: blockIdx.y = kz takes k in
// [kz kc, (kz + 1) kc) and writes its partial product to C + kz c_stride.
template <typename TOut, bool SYM = false>
__global__ __launch_bounds__(THREADS2, 1) void gemm_bf16x3_nt_256_kernel(
    int M, int N, int K, float alpha, const uint16_t* __restrict__ Ahi,
    const uint16_t* __restrict__ Alo, long long lda, const uint16_t* __restrict__ Bhi,
    const uint16_t* __restrict__ Blo, long long ldb, TOut* __restrict__ C, long long ldc,
    int tiles_m, int tiles_n, afq::sym_dst sd = afq::sym_dst{0, 0, 0}, int kc = 0, long long c_stride = 0,
    int lower_b = 0) {
    if (gridDim.y > 1) {
        const int k_begin = blockIdx.y * kc;
        K = min(K - k_begin, kc);
        Ahi += k_begin;
        Alo += k_begin;
        Bhi += k_begin;
        Blo += k_begin;
        C += blockIdx.y * c_stride;
    }
    // [stage][plane: hi, lo][256 rows x 64 bytes]: A two k-steps deep, B three
    // (the CU's 160 KB of LDS), so the B movers (waves 4..7) issue two k-steps
    // ahead and the A movers (waves 0..3) one
    __shared__ __attribute__((aligned(1024))) uint4 SA[2][2][BM2 * 4];
    __shared__ __attribute__((aligned(1024))) uint4 SB[3][2][BN2 * 4];

    const int total = tiles_m * tiles_n;
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int per_group = GROUP_M2 * tiles_n;
    const int group = t / per_group;
    const int first_m = group * GROUP_M2;
    const int gsize = min(tiles_m - first_m, GROUP_M2);
    const int in_group = t - group * per_group;
    const int m0 = (first_m + in_group % gsize) * BM2;
    const int n0 = (lower_b ? afq::zigzag(in_group / gsize, tiles_n) : in_group / gsize) * BN2;
    if (lower_b) K = min(K, n0 + BN2);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // wave-uniform (scalar): the descriptors and LDS bases below must be
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave & 3;   // 64-row slice of the tile
    const int wn = wave >> 2;  // 128-column slice

    // copy assignment: wave w moves plane w / 2 (A hi, A lo, B hi, B lo), 1 KB
    // blocks 8 (w % 2) + i, i < 8 (rows 16 blk ..); lane L fills slot L of a
    // block: row L / 4, position L % 4, i.e. chunk (L % 4) ^ ((L / 8) & 3)
    const int plane = wave >> 1;
    const uint16_t* src = plane == 0 ? Ahi + (long long)m0 * lda
                        : plane == 1 ? Alo + (long long)m0 * lda
                        : plane == 2 ? Bhi + (long long)n0 * ldb
                                     : Blo + (long long)n0 * ldb;
    const long long ld = plane < 2 ? lda : ldb;
    const int rows_left = max(plane < 2 ? M - m0 : N - n0, 0);
    const rsrc_t R = make_rsrc(src, (unsigned long long)rows_left * ld * 2);
    const int r_loc = lane >> 2;
    const int chunk = (lane & 3) ^ ((lane >> 3) & 3);
    const bool deep = plane >= 2;
    auto issue_stage = [&](int kt) {
        const int k = kt * 32 + chunk * 8;
        uint4* dst = deep ? &SB[kt % 3][plane - 2][0] : &SA[kt & 1][plane][0];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int blk = (wave & 1) * 8 + i;
            const int row = blk * 16 + r_loc;
            // k >= K sets the top bit: the offset lands outside the descriptor
            const unsigned past = (unsigned)((K - 1 - k) >> 31) & kOutside;
            const unsigned off = (unsigned)(((long long)row * ld + k) * 2) | past;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(R, (lds_ptr_t)(dst + blk * 64), 16, off, 0, 0, 0);
        }
    };

    f32x4 acc[4][8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // fragment read: row lr of a 16-row tile, chunk q at its swizzled position
    const int lr = lane & 15;
    const int q = lane >> 4;
    const int lane_slot = lr * 4 + (q ^ ((lr >> 1) & 3));

    const int nk = afq::ceil_div(K, 32);
    issue_stage(0);
    if (deep && nk > 1) {
        issue_stage(1);
        asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
    } else {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int ka = deep ? kt + 2 : kt + 1;  // the k-step this wave issues now
        const bool ahead = ka < nk;
        if (ahead) issue_stage(ka);
        const uint4* A0 = SA[kt & 1][0];
        const uint4* A1 = SA[kt & 1][1];
        const uint4* B0 = SB[kt % 3][0];
        const uint4* B1 = SB[kt % 3][1];
        bf16x8 ah[4], al[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int slot = (wm * 64 + i * 16) * 4 + lane_slot;
            ah[i] = __builtin_bit_cast(bf16x8, A0[slot]);
            al[i] = __builtin_bit_cast(bf16x8, A1[slot]);
        }
#pragma unroll
        for (int jh = 0; jh < 2; ++jh) {
            bf16x8 bh[4], bl[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int slot = (wn * 128 + (jh * 4 + j) * 16) * 4 + lane_slot;
                bh[j] = __builtin_bit_cast(bf16x8, B0[slot]);
                bl[j] = __builtin_bit_cast(bf16x8, B1[slot]);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    f32x4& c = acc[i][jh * 4 + j];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], c, 0, 0, 0);
                }
        }
        // k-step kt + 1 has landed in LDS and every wave is done with this one
        // (a B mover keeps its eight loads of k-step kt + 2 in flight)
        if (deep && ahead)
            asm volatile("s_waitcnt vmcnt(8)" ::: "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    }

    if constexpr (SYM) {
        // The 256-column tile is one block (bi, bj).  It leaves through LDS in
        // four chunks of 64 rows, so that every store writes whole 64-byte row
        // segments: the 16 entries of a block row at (16 bi + pi, 16 bj ..) and,
        // off the diagonal, the 16 entries of a block column at
        // (16 bj + qi, 16 bi ..).  Chunk layout: row c at c * RP floats, block
        // row pi at pi * PP (both padded: the fills and the column reads hit 64
        // distinct banks).
        constexpr int RP = 324, PP = 20;
        float* L = reinterpret_cast<float*>(&SB[0][0][0]);  // 64 x RP floats fit in B's 96 KB
        float* Cf = reinterpret_cast<float*>(C);
        int bi, bj;
        afq::sym_block(n0 >> 8, sd.nb, bi, bj);
        const long long ldf = (long long)sd.n * sd.ldv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
#pragma unroll
            for (int j = 0; j < 8; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) L[(wm * 16 + q * 4 + r) * RP + (wn * 8 + j) * PP + lr] = alpha * acc[i][j][r];
            __syncthreads();
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                const int u = tid + s * THREADS2;
                const int c = u >> 6;        // chunk row: wm = c / 16, 4 q + r = c % 16
                const int a = (u >> 2) & 15;  // block row (direct) / block column (mirror)
                const int b = u & 3;          // 4-entry group of the 64-byte segment
                const int row = m0 + (c >> 4) * 64 + i * 16 + (c & 15);
                if (row >= M) continue;
                float* dst = Cf + row * ldf;
                const int p = bi * 16 + a, qd = bj * 16 + 4 * b;
                if (p < sd.n && qd < sd.ldv)
                    *reinterpret_cast<float4*>(dst + (long long)p * sd.ldv + qd) =
                        *reinterpret_cast<const float4*>(&L[c * RP + a * PP + 4 * b]);
                const int qm = bj * 16 + a;
                if (bi != bj && qm < sd.n) {
                    const float* col = &L[c * RP + 4 * b * PP + a];
                    *reinterpret_cast<float4*>(dst + (long long)qm * sd.ldv + bi * 16 + 4 * b) =
                        make_float4(col[0], col[PP], col[2 * PP], col[3 * PP]);
                }
            }
            __syncthreads();
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int col = n0 + wn * 128 + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + q * 4 + r;
                if (row < M && col < N) C[(long long)row * ldc + col] = (TOut)(alpha * acc[i][j][r]);
            }
        }
}

// Grouped LDS-DMA GEMM (tile shapes 1 and 2: 128 x 256 and 256 x 256):
// workgroup -> (problem, tile) by the problems' tile_start prefix sums after
// the XCD-contiguous remap; the symmetric VHS store as in the 128 x 128
// grouped kernel (the 256-column tile is one block of V).
template <int WM>
__global__ __launch_bounds__(128 * WM, 1) void gemm_bf16x3_grouped_dma_kernel(
    int count, int total, const afq_gemm_problem_bf16x3* __restrict__ P) {
    using D = dma_shape<WM>;
    __shared__ __attribute__((aligned(1024))) uint4 S[D::SIZE];
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int pi = afq::find_group(count, t, [&](int i) { return P[i].tile_start; });
    const afq_gemm_problem_bf16x3& p = P[pi];
    const int tiles_m = afq::ceil_div(p.M, D::BM);
    const int local = t - p.tile_start;
    const int m0 = (local % tiles_m) * D::BM;
    const int nn = local / tiles_m;
    const int n0 = (p.b_lower ? afq::zigzag(nn, afq::ceil_div(p.N, 256)) : nn) * 256;
    const int K = p.b_lower ? min(p.K, n0 + 256) : p.K;
    f32x4 acc[4][8];
    dma_mainloop<WM>(p.M, p.N, K, p.A_hi, p.A_lo, p.lda, p.B_hi, p.B_lo, p.ldb, m0, n0, S, acc);
    if (p.sym_n > 0) {
        const afq::sym_dst sd{p.sym_n, p.sym_ldv, afq::ceil_div(p.sym_n, 16)};
        int bi, bj;
        afq::sym_block(n0 >> 8, sd.nb, bi, bj);
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const int wm = wave % WM, wn = wave / WM;
        const long long ldf = (long long)sd.n * sd.ldv;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
                if (row >= p.M) continue;
                float v[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) v[j] = p.alpha * acc[i][j][r];
                afq::sym_store<8>((float*)p.C + row * ldf, sd, bi, bj, wn * 8, lane & 15, v);
            }
    } else if (p.c_fp64 == 2) {
        dma_store<WM>(p.M, p.N, p.alpha, (afq::bf16_pair*)p.C, p.ldc, m0, n0, acc);
    } else if (p.c_fp64) {
        dma_store<WM>(p.M, p.N, p.alpha, (double*)p.C, p.ldc, m0, n0, acc);
    } else {
        dma_store<WM>(p.M, p.N, p.alpha, (float*)p.C, p.ldc, m0, n0, acc);
    }
}

// C = alpha * sum_kz part[kz] (kz in order), the split-K partials of M x N
__global__ void splitk_reduce_kernel(int M, int N, int splits, const float* __restrict__ part, long long stride,
                                     float alpha, void* __restrict__ C, long long ldc, int c_fp64) {
    const long long total = (long long)M * N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int z = 0; z < splits; ++z) acc += (double)part[z * stride + e];
        const long long r = e / N, c = e - r * N;
        if (c_fp64 == 2)
            ((afq::bf16_pair*)C)[r * ldc + c] = afq::bf16_pair((float)((double)alpha * acc));
        else if (c_fp64)
            ((double*)C)[r * ldc + c] = (double)alpha * acc;
        else
            ((float*)C)[r * ldc + c] = (float)((double)alpha * acc);
    }
}

__global__ void split_bf16_kernel(int rows, int cols, const double* __restrict__ src,
                                  long long ld_src, uint16_t* __restrict__ hi,
                                  uint16_t* __restrict__ lo, long long ld_dst) {
    const long long total = (long long)rows * ld_dst;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e / ld_dst;
        const long long c = e - r * ld_dst;
        const double x = c < cols ? src[r * ld_src + c] : 0.0;
        afq::split_bf16(x, hi[e], lo[e]);
    }
}

// Per problem (blockIdx.y): rows x cols of src (leading dimension ld_src) ->
// hi / lo planes of rows x ld_dst, columns cols..ld_dst-1 zero.
__global__ void split_bf16_grouped_kernel(const afq_split_problem* __restrict__ P) {
    const afq_split_problem& p = P[blockIdx.y];
    const long long total = (long long)p.rows * p.ld_dst;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e / p.ld_dst;
        const long long c = e - r * p.ld_dst;
        const double x = c < p.cols ? p.src[r * p.ld_src + c] : 0.0;
        afq::split_bf16(x, p.hi[e], p.lo[e]);
    }
}

__global__ void convert_f32_kernel(int rows, int cols, const double* __restrict__ src,
                                   long long ld_src, float* __restrict__ dst, long long ld_dst) {
    const long long total = (long long)rows * ld_dst;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e / ld_dst;
        const long long c = e - r * ld_dst;
        dst[e] = c < cols ? (float)src[r * ld_src + c] : 0.0f;
    }
}

unsigned grid_for(long long total) {
    long long b = (total + 255) / 256;
    return (unsigned)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

namespace {

// afq_gemm_bf16x3_nt and afq_gemm_bf16x3_nt_lower: 256 x 256 tiles once they
// give at least two workgroups per CU (one with a lower-triangular B), else
// 128 x 128 (the same sums in either tiling)
int bf16x3_nt(const char* name, int M, int N, int K, float alpha, const unsigned short* A_hi,
              const unsigned short* A_lo, long long lda, const unsigned short* B_hi, const unsigned short* B_lo,
              long long ldb, void* C, long long ldc, int c_fp64, int lower_b, hipStream_t st) {
    AFQ_REQUIRE(M >= 0 && N >= 0 && K >= 0, "%s: negative size", name);
    if (M == 0 || N == 0) return 0;
    AFQ_REQUIRE(A_hi && A_lo && B_hi && B_lo && C, "%s: null pointer", name);
    AFQ_REQUIRE(K % 8 == 0 && lda % 8 == 0 && ldb % 8 == 0, "%s: K=%d, lda and ldb must be multiples of 8", name, K);
    AFQ_REQUIRE(lda >= K && ldb >= K && ldc >= N, "%s: leading dimension too small", name);
    AFQ_REQUIRE(c_fp64 >= 0 && c_fp64 <= 2, "%s: c_fp64=%d not in 0..2", name, c_fp64);
    const afq::sym_dst none{0, 0, 0};
    const int tiles_m2 = afq::ceil_div(M, BM2);
    const int tiles_n2 = afq::ceil_div(N, BN2);
    const long long total2 = (long long)tiles_m2 * tiles_n2;
    // a lower-triangular B halves and evens out the work per tile: one round of
    // 256 x 256 tiles already fills the CUs
    const long long min2 = lower_b ? afq::cu_count() : 512;
    if (total2 >= min2 && total2 < (1LL << 31)) {
        const dim3 g((unsigned)total2);
        if (c_fp64 == 2)
            hipLaunchKernelGGL(gemm_bf16x3_nt_256_kernel<afq::bf16_pair>, g, dim3(THREADS2), 0, st, M, N, K, alpha,
                               A_hi, A_lo, lda, B_hi, B_lo, ldb, (afq::bf16_pair*)C, ldc, tiles_m2, tiles_n2, none, 0,
                               0LL, lower_b);
        else if (c_fp64)
            hipLaunchKernelGGL(gemm_bf16x3_nt_256_kernel<double>, g, dim3(THREADS2), 0, st, M, N, K, alpha, A_hi,
                               A_lo, lda, B_hi, B_lo, ldb, (double*)C, ldc, tiles_m2, tiles_n2, none, 0, 0LL,
                               lower_b);
        else
            hipLaunchKernelGGL(gemm_bf16x3_nt_256_kernel<float>, g, dim3(THREADS2), 0, st, M, N, K, alpha, A_hi,
                               A_lo, lda, B_hi, B_lo, ldb, (float*)C, ldc, tiles_m2, tiles_n2, none, 0, 0LL,
                               lower_b);
        return afq::check_launch(name);
    }
    const int tiles_m = afq::ceil_div(M, BM);
    const int tiles_n = afq::ceil_div(N, BN);
    const long long total = (long long)tiles_m * tiles_n;
    AFQ_REQUIRE(total < (1LL << 31), "%s: problem too large", name);
    const dim3 g((unsigned)total);
    if (c_fp64 == 2)
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<afq::bf16_pair>, g, dim3(THREADS), 0, st, M, N, K, alpha, A_hi, A_lo,
                           lda, B_hi, B_lo, ldb, (afq::bf16_pair*)C, ldc, tiles_m, tiles_n, none, 0, 0LL, lower_b);
    else if (c_fp64)
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<double>, g, dim3(THREADS), 0, st, M, N, K, alpha, A_hi, A_lo, lda,
                           B_hi, B_lo, ldb, (double*)C, ldc, tiles_m, tiles_n, none, 0, 0LL, lower_b);
    else
        hipLaunchKernelGGL(gemm_bf16x3_nt_kernel<float>, g, dim3(THREADS), 0, st, M, N, K, alpha, A_hi, A_lo, lda,
                           B_hi, B_lo, ldb, (float*)C, ldc, tiles_m, tiles_n, none, 0, 0LL, lower_b);
    return afq::check_launch(name);
}

}  // namespace

extern "C" int afq_gemm_bf16x3_nt(int M, int N, int K, float alpha, const unsigned short* A_hi,
                                  const unsigned short* A_lo, long long lda,
                                  const unsigned short* B_hi, const unsigned short* B_lo,
                                  long long ldb, void* C, long long ldc, int c_fp64, void* stream) {
    return bf16x3_nt("afq_gemm_bf16x3_nt", M, N, K, alpha, A_hi, A_lo, lda, B_hi, B_lo, ldb, C, ldc, c_fp64, 0,
                     (hipStream_t)stream);
}

extern "C" int afq_gemm_bf16x3_nt_lower(int M, int N, int K, float alpha, const unsigned short* A_hi,
                                        const unsigned short* A_lo, long long lda, const unsigned short* B_hi,
                                        const unsigned short* B_lo, long long ldb, void* C, long long ldc,
                                        int c_fp64, void* stream) {
    return bf16x3_nt("afq_gemm_bf16x3_nt_lower", M, N, K, alpha, A_hi, A_lo, lda, B_hi, B_lo, ldb, C, ldc, c_fp64,
                     1, (hipStream_t)stream);
}

extern "C" int afq_gemm_bf16x3_nt_splitk(int M, int N, int K, float alpha, const unsigned short* A_hi,
                                         const unsigned short* A_lo, long long lda, const unsigned short* B_hi,
                                         const unsigned short* B_lo, long long ldb, void* C, long long ldc,
                                         int c_fp64, int splits, float* workspace, void* stream) {
    AFQ_REQUIRE(M >= 0 && N >= 0 && K >= 0, "afq_gemm_bf16x3_nt_splitk: negative size");
    if (M == 0 || N == 0) return 0;
    AFQ_REQUIRE(splits >= 1 && splits <= 16, "afq_gemm_bf16x3_nt_splitk: splits=%d not in 1..16", splits);
    AFQ_REQUIRE(c_fp64 >= 0 && c_fp64 <= 2, "afq_gemm_bf16x3_nt_splitk: c_fp64=%d not in 0..2", c_fp64);
    AFQ_REQUIRE(A_hi && A_lo && B_hi && B_lo && C && workspace, "afq_gemm_bf16x3_nt_splitk: null pointer");
    AFQ_REQUIRE(K % 8 == 0 && lda % 8 == 0 && ldb % 8 == 0,
                "afq_gemm_bf16x3_nt_splitk: K=%d, lda and ldb must be multiples of 8", K);
    AFQ_REQUIRE(lda >= K && ldb >= K && ldc >= N, "afq_gemm_bf16x3_nt_splitk: leading dimension too small");
    const int tiles_m = afq::ceil_div(M, BM);
    const int tiles_n = afq::ceil_div(N, BN);
    AFQ_REQUIRE((long long)tiles_m * tiles_n < (1LL << 31), "afq_gemm_bf16x3_nt_splitk: problem too large");
    // K chunks of whole 32-wide k-steps; no empty chunk
    const int kc = afq::round_up(afq::ceil_div(K, splits), 32);
    const int s = K == 0 ? 1 : afq::ceil_div(K, kc);
    const long long stride = (long long)M * N;
    hipStream_t st = (hipStream_t)stream;
    // 256 x 256 tiles (one workgroup per CU) when their chunks fill the CUs in
    // rounds at least three quarters full; each chunk's partial product is the
    // same in either tiling (same k order, same three products per k-step)
    const long long tiles2 = (long long)afq::ceil_div(M, BM2) * afq::ceil_div(N, BN2);
    const long long wg2 = tiles2 * s;
    const long long cus = afq::cu_count();
    const long long rounds = (wg2 + cus - 1) / cus;
    if (tiles2 < (1LL << 31) && wg2 >= cus && 4 * wg2 >= 3 * rounds * cus) {
        hipLaunchKernelGGL((gemm_bf16x3_nt_256_kernel<float, false>), dim3((unsigned)tiles2, (unsigned)s),
                           dim3(THREADS2), 0, st, M, N, K, 1.0f, A_hi, A_lo, lda, B_hi, B_lo, ldb, workspace,
                           (long long)N, afq::ceil_div(M, BM2), afq::ceil_div(N, BN2), afq::sym_dst{0, 0, 0}, kc,
                           stride);
    } else {
        hipLaunchKernelGGL((gemm_bf16x3_nt_kernel<float, false>), dim3((unsigned)(tiles_m * tiles_n), (unsigned)s),
                           dim3(THREADS), 0, st, M, N, K, 1.0f, A_hi, A_lo, lda, B_hi, B_lo, ldb, workspace,
                           (long long)N, tiles_m, tiles_n, afq::sym_dst{0, 0, 0}, kc, stride);
    }
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for(stride)), dim3(256), 0, st, M, N, s, workspace, stride,
                       alpha, C, ldc, c_fp64);
    return afq::check_launch("afq_gemm_bf16x3_nt_splitk");
}

extern "C" int afq_vhs_sym(int rows, int n, int ldv, int k, int arith, const void* a_hi, const void* a_lo,
                           long long lda, const void* b_hi, const void* b_lo, long long ldb, void* vhs,
                           void* stream) {
    AFQ_REQUIRE(rows > 0 && n > 0 && k > 0, "afq_vhs_sym: bad sizes rows=%d n=%d k=%d", rows, n, k);
    AFQ_REQUIRE(ldv >= n && ldv % 4 == 0, "afq_vhs_sym: ldv=%d must be >= n=%d and a multiple of 4", ldv, n);
    AFQ_REQUIRE(arith >= 0 && arith <= 2, "afq_vhs_sym: arith=%d not in 0..2", arith);
    AFQ_REQUIRE(a_hi && b_hi && vhs && (arith != 2 || (a_lo && b_lo)), "afq_vhs_sym: null pointer");
    AFQ_REQUIRE(((uintptr_t)vhs & 15) == 0, "afq_vhs_sym: vhs must be 16-byte aligned");
    AFQ_REQUIRE(lda >= k && ldb >= k, "afq_vhs_sym: leading dimension too small");
    const int nb = afq::ceil_div(n, 16);
    AFQ_REQUIRE((long long)nb * (nb + 1) / 2 * 256 < (1LL << 31), "afq_vhs_sym: n=%d too large", n);
    const afq::sym_dst d{n, ldv, nb};
    hipStream_t st = (hipStream_t)stream;
    if (arith == 0)
        return afq::gemm_nt_sym<double>("afq_vhs_sym", rows, k, (const double*)a_hi, lda, (const double*)b_hi, ldb,
                                        d, (double*)vhs, st);
    if (arith == 1)
        return afq::gemm_nt_sym<float>("afq_vhs_sym", rows, k, (const float*)a_hi, lda, (const float*)b_hi, ldb,
                                       d, (float*)vhs, st);
    AFQ_REQUIRE(k % 8 == 0 && lda % 8 == 0 && ldb % 8 == 0,
                "afq_vhs_sym: bf16x3 needs k=%d, lda and ldb multiples of 8", k);
    const int N = nb * (nb + 1) / 2 * 256;
    const uint16_t *ah = (const uint16_t*)a_hi, *al = (const uint16_t*)a_lo;
    const uint16_t *bh = (const uint16_t*)b_hi, *bl = (const uint16_t*)b_lo;
    const int tiles_m2 = afq::ceil_div(rows, BM2);
    const int tiles_n2 = N / BN2;
    if ((long long)tiles_m2 * tiles_n2 >= 512) {
        hipLaunchKernelGGL((gemm_bf16x3_nt_256_kernel<float, true>), dim3((unsigned)(tiles_m2 * tiles_n2)),
                           dim3(THREADS2), 0, st, rows, N, k, 1.0f, ah, al, lda, bh, bl, ldb, (float*)vhs, 0LL,
                           tiles_m2, tiles_n2, d);
    } else {
        const int tiles_m = afq::ceil_div(rows, BM);
        const int tiles_n = N / BN;
        hipLaunchKernelGGL((gemm_bf16x3_nt_kernel<float, true>), dim3((unsigned)(tiles_m * tiles_n)),
                           dim3(THREADS), 0, st, rows, N, k, 1.0f, ah, al, lda, bh, bl, ldb, (float*)vhs, 0LL,
                           tiles_m, tiles_n, d);
    }
    return afq::check_launch("afq_vhs_sym");
}

extern "C" int afq_split_bf16(int rows, int cols, const double* src, long long ld_src,
                              unsigned short* hi, unsigned short* lo, long long ld_dst,
                              void* stream) {
    AFQ_REQUIRE(rows >= 0 && cols >= 0 && ld_src >= cols && ld_dst >= cols,
                "afq_split_bf16: bad sizes rows=%d cols=%d", rows, cols);
    if (rows == 0 || ld_dst == 0) return 0;
    AFQ_REQUIRE(src && hi && lo, "afq_split_bf16: null pointer");
    hipLaunchKernelGGL(split_bf16_kernel, dim3(grid_for((long long)rows * ld_dst)), dim3(256), 0,
                       (hipStream_t)stream, rows, cols, src, ld_src, hi, lo, ld_dst);
    return afq::check_launch("afq_split_bf16");
}

extern "C" int afq_convert_f32(int rows, int cols, const double* src, long long ld_src, float* dst,
                               long long ld_dst, void* stream) {
    AFQ_REQUIRE(rows >= 0 && cols >= 0 && ld_src >= cols && ld_dst >= cols,
                "afq_convert_f32: bad sizes rows=%d cols=%d", rows, cols);
    if (rows == 0 || ld_dst == 0) return 0;
    AFQ_REQUIRE(src && dst, "afq_convert_f32: null pointer");
    hipLaunchKernelGGL(convert_f32_kernel, dim3(grid_for((long long)rows * ld_dst)), dim3(256), 0,
                       (hipStream_t)stream, rows, cols, src, ld_src, dst, ld_dst);
    return afq::check_launch("afq_convert_f32");
}

extern "C" int afq_gemm_bf16x3_tile_shape(int shape, int* bm, int* bn) {
    AFQ_REQUIRE(shape >= 0 && shape <= 2, "afq_gemm_bf16x3_tile_shape: shape=%d not in 0..2", shape);
    AFQ_REQUIRE(bm && bn, "afq_gemm_bf16x3_tile_shape: null pointer");
    *bm = shape == 0 ? BM : shape == 1 ? dma_shape<2>::BM : dma_shape<4>::BM;
    *bn = shape == 0 ? BN : 256;
    return 0;
}

extern "C" int afq_gemm_bf16x3_nt_grouped(int count, const afq_gemm_problem_bf16x3* problems, int total_tiles,
                                          int shape, void* stream) {
    AFQ_REQUIRE(count > 0 && total_tiles > 0, "afq_gemm_bf16x3_nt_grouped: count=%d total_tiles=%d", count,
                total_tiles);
    AFQ_REQUIRE(shape >= 0 && shape <= 2, "afq_gemm_bf16x3_nt_grouped: shape=%d not in 0..2", shape);
    AFQ_REQUIRE(problems, "afq_gemm_bf16x3_nt_grouped: null pointer");
    hipStream_t st = (hipStream_t)stream;
    if (shape == 0)
        hipLaunchKernelGGL(gemm_bf16x3_grouped_kernel, dim3((unsigned)total_tiles), dim3(THREADS), 0, st, count,
                           total_tiles, problems);
    else if (shape == 1)
        hipLaunchKernelGGL(gemm_bf16x3_grouped_dma_kernel<2>, dim3((unsigned)total_tiles),
                           dim3(dma_shape<2>::THREADS), 0, st, count, total_tiles, problems);
    else
        hipLaunchKernelGGL(gemm_bf16x3_grouped_dma_kernel<4>, dim3((unsigned)total_tiles),
                           dim3(dma_shape<4>::THREADS), 0, st, count, total_tiles, problems);
    return afq::check_launch("afq_gemm_bf16x3_nt_grouped");
}

extern "C" int afq_split_bf16_grouped(int count, const afq_split_problem* problems, long long max_elems,
                                      void* stream) {
    AFQ_REQUIRE(count > 0 && count <= 65535, "afq_split_bf16_grouped: count=%d not in 1..65535", count);
    AFQ_REQUIRE(max_elems > 0, "afq_split_bf16_grouped: max_elems=%lld", max_elems);
    AFQ_REQUIRE(problems, "afq_split_bf16_grouped: null pointer");
    long long b = (max_elems + 255) / 256;
    if (b > 2048) b = 2048;
    hipLaunchKernelGGL(split_bf16_grouped_kernel, dim3((unsigned)b, (unsigned)count), dim3(256), 0,
                       (hipStream_t)stream, problems);
    return afq::check_launch("afq_split_bf16_grouped");
}
