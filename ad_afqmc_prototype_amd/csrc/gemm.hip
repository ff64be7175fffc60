// GEMMs  C = alpha * A * B^T + beta * C  on gfx950 matrix cores:
//   fp64   v_mfma_f64_16x16x4_f64           (plain fp64 step)
//   fp32   v_mfma_f32_16x16x4_f32           (mixed step, fp32 contractions)
//   bf16x3 3 x v_mfma_f32_16x16x32_bf16    (mixed step, split-bf16 contractions:
//          A B ~ Ahi Bhi + Ahi Blo + Alo Bhi, fp32 accumulation; Henry, Tang &
//          Heinecke, ARITH 2019)
//
// Both operands are K-contiguous (row-major A: M x K, B: N x K), which is the
// natural layout of the force-bias contraction (greens rows . rotated Cholesky
// rows) and of the VHS contraction (shifted-field rows . L^T rows).
//
// Tiling: 128 x 128 x 16 per 256-thread workgroup (4 waves in 2 x 2, each wave
// 64 x 64 = 4 x 4 MFMA tiles), operands staged through double-buffered LDS
// (73.7 KB, two workgroups per CU).  Each lane reads four consecutive k of one
// row from LDS and feeds them to four MFMAs; A and B use the same k
// permutation, so the sum over the 16 k of a block is complete.  Workgroups are
// remapped so that each XCD owns a contiguous range of tiles, and tiles are
// ordered in groups of 8 M-tiles so that concurrently running workgroups share
// A and B panels in L2.
#include <algorithm>

#include "afqmc_hip.h"
#include "common.h"

namespace {

constexpr int BK = 16;
constexpr int PITCH = BK + 2;  // 144-byte rows: conflict-free 16-byte LDS reads
constexpr int THREADS = 256;
constexpr int GROUP_M = 8;

// One workgroup computes a BM x BN tile of C, BM = 32 TM, BN = 32 TN: four
// waves in 2 x 2, each TM x TN MFMA tiles of 16 x 16.
template <typename T, int TM, int TN>
struct tile_cfg {
    static constexpr int BM = 32 * TM;
    static constexpr int BN = 32 * TN;
    static constexpr int LA = BM * BK / THREADS;  // A elements loaded per thread per k-block
    static constexpr int LB = BN * BK / THREADS;
    using acc_t = typename afq::mfma16<T>::acc;
};

// acc = A[m0:m0+BM, :] . B[n0:n0+BN, :]^T over all of K, through
// double-buffered LDS (As: 2 * BM * PITCH, Bs: 2 * BN * PITCH elements).
template <typename T, int TM, int TN>
__device__ __forceinline__ void gemm_nt_tile(int M, int N, int K, const T* __restrict__ A,
                                             long long lda, const T* __restrict__ B, long long ldb,
                                             int m0, int n0, T* As, T* Bs,
                                             typename tile_cfg<T, TM, TN>::acc_t (&acc)[TM][TN]) {
    using C = tile_cfg<T, TM, TN>;
    using MF = afq::mfma16<T>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1;
    const int wn = wave & 1;

    T ra[C::LA], rb[C::LB];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int i = 0; i < C::LA; ++i) {
            const int e = tid + THREADS * i;
            const int r = e >> 4;
            const int gk = k0 + (e & 15);
            const int gm = m0 + r;
            ra[i] = (gm < M && gk < K) ? A[(long long)gm * lda + gk] : T(0);
        }
#pragma unroll
        for (int i = 0; i < C::LB; ++i) {
            const int e = tid + THREADS * i;
            const int r = e >> 4;
            const int gk = k0 + (e & 15);
            const int gn = n0 + r;
            rb[i] = (gn < N && gk < K) ? B[(long long)gn * ldb + gk] : T(0);
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int i = 0; i < C::LA; ++i) {
            const int e = tid + THREADS * i;
            As[buf * C::BM * PITCH + (e >> 4) * PITCH + (e & 15)] = ra[i];
        }
#pragma unroll
        for (int i = 0; i < C::LB; ++i) {
            const int e = tid + THREADS * i;
            Bs[buf * C::BN * PITCH + (e >> 4) * PITCH + (e & 15)] = rb[i];
        }
    };

#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = MF::zero();

    const int nk = afq::ceil_div(K, BK);
    load_tile(0);
    store_tile(0);
    __syncthreads();

    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tile((kt + 1) * BK);
        const T* as = As + buf * C::BM * PITCH;
        const T* bs = Bs + buf * C::BN * PITCH;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            T a[TM], b[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = as[(wm * 16 * TM + i * 16 + lr) * PITCH + lk + kk];
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = bs[(wn * 16 * TN + j * 16 + lr) * PITCH + lk + kk];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] = MF::fma(a[i], b[j], acc[i][j]);
        }
        if (kt + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }
}

// C[m0.., n0..] = alpha acc + beta C
template <typename T, typename TOut, int TM, int TN>
__device__ __forceinline__ void gemm_store(int M, int N, T alpha, T beta, TOut* __restrict__ C,
                                           long long ldc, int m0, int n0,
                                           const typename tile_cfg<T, TM, TN>::acc_t (&acc)[TM][TN]) {
    using MF = afq::mfma16<T>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave >> 1;
    const int wn = wave & 1;
    const int lr = lane & 15;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = n0 + wn * 16 * TN + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 16 * TM + i * 16 + MF::row(lane, r);
                if (row < M && col < N) {
                    TOut* cp = C + (long long)row * ldc + col;
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * (T)*cp;
                    *cp = (TOut)v;
                }
            }
        }
    }
}

// The store of a tile whose columns are symmetric-VHS block entries
// (afq::sym_store): a 128-column tile lies inside one 256-column block, and a
// wave's TN column tiles are TN consecutive block rows pi.
template <typename T, int TM, int TN>
__device__ __forceinline__ void gemm_store_sym(int M, T alpha, T* __restrict__ C, const afq::sym_dst& d,
                                               int m0, int n0,
                                               const typename tile_cfg<T, TM, TN>::acc_t (&acc)[TM][TN]) {
    using MF = afq::mfma16<T>;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave >> 1;
    const int wn = wave & 1;
    int bi, bj;
    afq::sym_block(n0 >> 8, d.nb, bi, bj);
    const int pi0 = ((n0 & 255) >> 4) + wn * TN;
    const long long ldc = (long long)d.n * d.ldv;
#pragma unroll
    for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = m0 + wm * 16 * TM + i * 16 + MF::row(lane, r);
            if (row >= M) continue;
            T v[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j) v[j] = alpha * acc[i][j][r];
            afq::sym_store<TN>(C + row * ldc, d, bi, bj, pi0, lane & 15, v);
        }
    }
}

// Origin (m0, n0) of tile blockIdx.x of a tiles_m x tiles_n grid: XCD-contiguous
// remap, then groups of GROUP_M tile rows walked column by column.
__device__ __forceinline__ void tile_origin(int tiles_m, int tiles_n, int bm, int bn, int& m0, int& n0) {
    const int total = tiles_m * tiles_n;
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int per_group = GROUP_M * tiles_n;
    const int group = t / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int in_group = t - group * per_group;
    m0 = (first_m + in_group % gsize) * bm;
    n0 = (in_group / gsize) * bn;
}

// lower_b: row j of B is zero past column j, so the tile of columns n0.. stops
// its k-loop at k = n0 + BN (the zeros inside that range are still read); the
// column tiles are then taken in afq::zigzag order, so that every XCD's
// contiguous share of tiles holds the same work
template <typename T, typename TOut, bool SYM = false>
__global__ __launch_bounds__(THREADS, 2) void gemm_nt_kernel(
    int M, int N, int K, T alpha, const T* __restrict__ A, long long lda,
    const T* __restrict__ B, long long ldb, T beta, TOut* __restrict__ C,
    long long ldc, int tiles_m, int tiles_n, afq::sym_dst sd = afq::sym_dst{0, 0, 0}, int lower_b = 0) {
    using CF = tile_cfg<T, 4, 4>;
    __shared__ T As[2 * CF::BM * PITCH];
    __shared__ T Bs[2 * CF::BN * PITCH];

    int m0, n0;
    tile_origin(tiles_m, tiles_n, CF::BM, CF::BN, m0, n0);
    if (lower_b) {
        n0 = afq::zigzag(n0 / CF::BN, tiles_n) * CF::BN;
        K = min(K, n0 + CF::BN);
    }

    typename CF::acc_t acc[4][4];
    gemm_nt_tile<T, 4, 4>(M, N, K, A, lda, B, ldb, m0, n0, As, Bs, acc);
    if constexpr (SYM)
        gemm_store_sym<T, 4, 4>(M, alpha, (T*)C, sd, m0, n0, acc);
    else
        gemm_store<T, TOut, 4, 4>(M, N, alpha, beta, C, ldc, m0, n0, acc);
}

// One K chunk of a split-K fp64 GEMM: chunk blockIdx.y covers k in
// [kz kc, min(K, (kz + 1) kc)) and leaves its partial product, unscaled, at
// part + kz M N (row-major, leading dimension N).
__global__ __launch_bounds__(THREADS, 2) void dgemm_nt_splitk_kernel(
    int M, int N, int K, int kc, const double* __restrict__ A, long long lda, const double* __restrict__ B,
    long long ldb, double* __restrict__ part, int tiles_m, int tiles_n) {
    using CF = tile_cfg<double, 4, 4>;
    __shared__ double As[2 * CF::BM * PITCH];
    __shared__ double Bs[2 * CF::BN * PITCH];
    int m0, n0;
    tile_origin(tiles_m, tiles_n, CF::BM, CF::BN, m0, n0);
    const int k0 = blockIdx.y * kc;
    typename CF::acc_t acc[4][4];
    gemm_nt_tile<double, 4, 4>(M, N, min(kc, K - k0), A + k0, lda, B + k0, ldb, m0, n0, As, Bs, acc);
    gemm_store<double, double, 4, 4>(M, N, 1.0, 0.0, part + (long long)blockIdx.y * M * N, N, m0, n0, acc);
}

// C = alpha * (part[0] + part[1] + ...), the chunks summed in order
__global__ void dsplitk_reduce_kernel(int M, int N, int splits, const double* __restrict__ part, double alpha,
                                      double* __restrict__ C, long long ldc) {
    const long long total = (long long)M * N;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        double acc = part[e];
        for (int z = 1; z < splits; ++z) acc += part[z * total + e];
        const long long r = e / N, c = e - r * N;
        C[r * ldc + c] = alpha * acc;
    }
}

// Grouped fp32 GEMM: workgroup -> (problem, tile) by the problems' tile_start
// prefix sums; the linear workgroup id is first made XCD-contiguous, so each
// XCD runs a contiguous range of tiles (of one or a few problems) and shares
// their panels in its L2.
template <int TM, int TN>
__global__ __launch_bounds__(THREADS, 2) void sgemm_grouped_kernel(int count, int total,
                                                                   const afq_gemm_problem* __restrict__ P) {
    using CF = tile_cfg<float, TM, TN>;
    __shared__ float As[2 * CF::BM * PITCH];
    __shared__ float Bs[2 * CF::BN * PITCH];
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int pi = afq::find_group(count, t, [&](int i) { return P[i].tile_start; });
    const afq_gemm_problem& p = P[pi];
    const int M = p.M, N = p.N;
    const int tiles_m = afq::ceil_div(M, CF::BM);
    const int local = t - p.tile_start;
    const int m0 = (local % tiles_m) * CF::BM;
    const int nn = local / tiles_m;
    // b_lower: column tiles in zigzag order, each k-loop stopping at its last column
    const int n0 = (p.b_lower ? afq::zigzag(nn, afq::ceil_div(N, CF::BN)) : nn) * CF::BN;
    const int K = p.b_lower ? min(p.K, n0 + CF::BN) : p.K;
    typename CF::acc_t acc[TM][TN];
    gemm_nt_tile<float, TM, TN>(M, N, K, p.A, p.lda, p.B, p.ldb, m0, n0, As, Bs, acc);
    if (p.sym_n > 0)
        gemm_store_sym<float, TM, TN>(M, p.alpha, (float*)p.C,
                                      afq::sym_dst{p.sym_n, p.sym_ldv, afq::ceil_div(p.sym_n, 16)}, m0, n0, acc);
    else if (p.c_fp64)
        gemm_store<float, double, TM, TN>(M, N, p.alpha, 0.0f, (double*)p.C, p.ldc, m0, n0, acc);
    else
        gemm_store<float, float, TM, TN>(M, N, p.alpha, 0.0f, (float*)p.C, p.ldc, m0, n0, acc);
}

constexpr int GROUPED_SHAPES[4][2] = {{4, 4}, {2, 4}, {4, 2}, {2, 2}};

}  // namespace

namespace afq {

template <typename T, typename TOut>
int gemm_nt(const char* name, int M, int N, int K, T alpha, const T* A, long long lda, const T* B,
            long long ldb, T beta, TOut* C, long long ldc, hipStream_t stream, int lower_b = 0) {
    AFQ_REQUIRE(M >= 0 && N >= 0 && K >= 0, "%s: negative size (M=%d N=%d K=%d)", name, M, N, K);
    if (M == 0 || N == 0) return 0;
    AFQ_REQUIRE(A && B && C, "%s: null pointer", name);
    AFQ_REQUIRE(lda >= K && ldb >= K && ldc >= N, "%s: leading dimension too small", name);
    const int tiles_m = ceil_div(M, tile_cfg<T, 4, 4>::BM);
    const int tiles_n = ceil_div(N, tile_cfg<T, 4, 4>::BN);
    const long long total = (long long)tiles_m * tiles_n;
    AFQ_REQUIRE(total < (1LL << 31), "%s: problem too large", name);
    hipLaunchKernelGGL((gemm_nt_kernel<T, TOut>), dim3((unsigned)total), dim3(THREADS), 0, stream, M, N, K,
                       alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n, sym_dst{0, 0, 0}, lower_b);
    return check_launch(name);
}

// symmetric-VHS GEMM (alpha = 1): M rows of A (M x K) against the upper-block
// rows of B (N = blocks * 256 rows), stored expanded into C (M x n x ldv)
template <typename T>
int gemm_nt_sym(const char* name, int M, int K, const T* A, long long lda, const T* B, long long ldb,
                sym_dst d, T* C, hipStream_t stream) {
    const int N = d.nb * (d.nb + 1) / 2 * 256;
    const int tiles_m = ceil_div(M, tile_cfg<T, 4, 4>::BM);
    const int tiles_n = ceil_div(N, tile_cfg<T, 4, 4>::BN);
    const long long total = (long long)tiles_m * tiles_n;
    AFQ_REQUIRE(total < (1LL << 31), "%s: problem too large", name);
    hipLaunchKernelGGL((gemm_nt_kernel<T, T, true>), dim3((unsigned)total), dim3(THREADS), 0, stream, M, N, K,
                       T(1), A, lda, B, ldb, T(0), C, 0LL, tiles_m, tiles_n, d);
    return check_launch(name);
}

template int gemm_nt_sym<double>(const char*, int, int, const double*, long long, const double*, long long,
                                 sym_dst, double*, hipStream_t);
template int gemm_nt_sym<float>(const char*, int, int, const float*, long long, const float*, long long,
                                sym_dst, float*, hipStream_t);

int dgemm_nt(int M, int N, int K, double alpha, const double* A, long long lda, const double* B,
             long long ldb, double beta, double* C, long long ldc, hipStream_t stream) {
    return gemm_nt<double, double>("afq_dgemm_nt", M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                                   stream);
}

}  // namespace afq

extern "C" int afq_sgemm_nt(int M, int N, int K, float alpha, const float* A, long long lda,
                            const float* B, long long ldb, float beta, void* C, long long ldc,
                            int c_fp64, void* stream) {
    if (c_fp64)
        return afq::gemm_nt<float, double>("afq_sgemm_nt", M, N, K, alpha, A, lda, B, ldb, beta,
                                           (double*)C, ldc, (hipStream_t)stream);
    return afq::gemm_nt<float, float>("afq_sgemm_nt", M, N, K, alpha, A, lda, B, ldb, beta,
                                      (float*)C, ldc, (hipStream_t)stream);
}

extern "C" int afq_dgemm_nt(int M, int N, int K, double alpha, const double* A, long long lda,
                            const double* B, long long ldb, double beta, double* C, long long ldc,
                            void* stream) {
    return afq::dgemm_nt(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, (hipStream_t)stream);
}

extern "C" int afq_dgemm_nt_lower(int M, int N, int K, double alpha, const double* A, long long lda,
                                  const double* B, long long ldb, double* C, long long ldc, void* stream) {
    return afq::gemm_nt<double, double>("afq_dgemm_nt_lower", M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc,
                                        (hipStream_t)stream, 1);
}

extern "C" int afq_sgemm_nt_lower(int M, int N, int K, float alpha, const float* A, long long lda,
                                  const float* B, long long ldb, void* C, long long ldc, int c_fp64,
                                  void* stream) {
    if (c_fp64)
        return afq::gemm_nt<float, double>("afq_sgemm_nt_lower", M, N, K, alpha, A, lda, B, ldb, 0.0f,
                                           (double*)C, ldc, (hipStream_t)stream, 1);
    return afq::gemm_nt<float, float>("afq_sgemm_nt_lower", M, N, K, alpha, A, lda, B, ldb, 0.0f, (float*)C,
                                      ldc, (hipStream_t)stream, 1);
}

extern "C" int afq_dgemm_nt_splitk(int M, int N, int K, double alpha, const double* A, long long lda,
                                   const double* B, long long ldb, double* C, long long ldc, int splits,
                                   double* workspace, void* stream) {
    AFQ_REQUIRE(M >= 0 && N >= 0 && K >= 0, "afq_dgemm_nt_splitk: negative size (M=%d N=%d K=%d)", M, N, K);
    if (M == 0 || N == 0) return 0;
    AFQ_REQUIRE(splits >= 1 && splits <= 16, "afq_dgemm_nt_splitk: splits=%d not in 1..16", splits);
    AFQ_REQUIRE(A && B && C && workspace, "afq_dgemm_nt_splitk: null pointer");
    AFQ_REQUIRE(lda >= K && ldb >= K && ldc >= N, "afq_dgemm_nt_splitk: leading dimension too small");
    using CF = tile_cfg<double, 4, 4>;
    const int tiles_m = afq::ceil_div(M, CF::BM);
    const int tiles_n = afq::ceil_div(N, CF::BN);
    AFQ_REQUIRE((long long)tiles_m * tiles_n < (1LL << 31), "afq_dgemm_nt_splitk: problem too large");
    // chunks of whole 16-deep k-blocks, none empty
    const int kc = K == 0 ? BK : afq::round_up(afq::ceil_div(K, splits), BK);
    const int s = K == 0 ? 1 : afq::ceil_div(K, kc);
    const long long total = (long long)M * N;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(dgemm_nt_splitk_kernel, dim3((unsigned)(tiles_m * tiles_n), (unsigned)s), dim3(THREADS), 0, st,
                       M, N, K, kc, A, lda, B, ldb, workspace, tiles_m, tiles_n);
    const long long blocks = std::min<long long>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(dsplitk_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, st, M, N, s, workspace, alpha, C,
                       ldc);
    return afq::check_launch("afq_dgemm_nt_splitk");
}

extern "C" int afq_force_bias(int nw, int naux, int k, double alpha, const double* greens,
                              const double* rot, double* fb, void* stream) {
    AFQ_REQUIRE(nw > 0 && naux > 0 && k > 0, "afq_force_bias: bad sizes nw=%d naux=%d k=%d", nw, naux, k);
    // rows of A: (plane, walker); rows of B: Cholesky vector g.
    return afq::dgemm_nt(2 * nw, naux, k, alpha, greens, k, rot, k, 0.0, fb, naux,
                         (hipStream_t)stream);
}

extern "C" int afq_vhs(int nw, int n, int ldv, int naux, const double* xbar, const double* chol_t,
                       double* vhs, void* stream) {
    AFQ_REQUIRE(nw > 0 && n > 0 && naux > 0, "afq_vhs: bad sizes nw=%d n=%d naux=%d", nw, n, naux);
    AFQ_REQUIRE(ldv >= n && ldv % 4 == 0, "afq_vhs: ldv=%d must be >= n=%d and a multiple of 4", ldv, n);
    return afq::dgemm_nt(2 * nw, n * ldv, naux, 1.0, xbar, naux, chol_t, naux, 0.0, vhs,
                         (long long)n * ldv, (hipStream_t)stream);
}

extern "C" int afq_gemm_tile_shape(int shape, int* bm, int* bn) {
    AFQ_REQUIRE(shape >= 0 && shape < 4, "afq_gemm_tile_shape: shape=%d not in 0..3", shape);
    AFQ_REQUIRE(bm && bn, "afq_gemm_tile_shape: null pointer");
    *bm = 32 * GROUPED_SHAPES[shape][0];
    *bn = 32 * GROUPED_SHAPES[shape][1];
    return 0;
}

extern "C" int afq_sgemm_nt_grouped(int count, const afq_gemm_problem* problems, int total_tiles,
                                    int shape, void* stream) {
    AFQ_REQUIRE(count > 0 && total_tiles > 0, "afq_sgemm_nt_grouped: count=%d total_tiles=%d",
                count, total_tiles);
    AFQ_REQUIRE(problems, "afq_sgemm_nt_grouped: null pointer");
    AFQ_REQUIRE(shape >= 0 && shape < 4, "afq_sgemm_nt_grouped: shape=%d not in 0..3", shape);
    hipStream_t st = (hipStream_t)stream;
    switch (shape) {
        case 0: hipLaunchKernelGGL((sgemm_grouped_kernel<4, 4>), dim3(total_tiles), dim3(THREADS), 0, st, count, total_tiles, problems); break;
        case 1: hipLaunchKernelGGL((sgemm_grouped_kernel<2, 4>), dim3(total_tiles), dim3(THREADS), 0, st, count, total_tiles, problems); break;
        case 2: hipLaunchKernelGGL((sgemm_grouped_kernel<4, 2>), dim3(total_tiles), dim3(THREADS), 0, st, count, total_tiles, problems); break;
        default: hipLaunchKernelGGL((sgemm_grouped_kernel<2, 2>), dim3(total_tiles), dim3(THREADS), 0, st, count, total_tiles, problems); break;
    }
    return afq::check_launch("afq_sgemm_nt_grouped");
}
