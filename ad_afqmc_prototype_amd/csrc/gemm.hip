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
#include "afqmc_hip.h"
#include "common.h"

namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int PITCH = BK + 2;  // 144-byte rows: conflict-free 16-byte LDS reads
constexpr int THREADS = 256;
constexpr int GROUP_M = 8;
constexpr int LOADS = BM * BK / THREADS;  // 8 elements of A and of B per thread

template <typename T, typename TOut>
__global__ __launch_bounds__(THREADS, 2) void gemm_nt_kernel(
    int M, int N, int K, T alpha, const T* __restrict__ A, long long lda,
    const T* __restrict__ B, long long ldb, T beta, TOut* __restrict__ C,
    long long ldc, int tiles_m, int tiles_n) {
    using MF = afq::mfma16<T>;
    __shared__ T As[2][BM * PITCH];
    __shared__ T Bs[2][BN * PITCH];

    const int total = tiles_m * tiles_n;
    const int t = afq::xcd_contiguous(blockIdx.x, total);
    const int per_group = GROUP_M * tiles_n;
    const int group = t / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(tiles_m - first_m, GROUP_M);
    const int in_group = t - group * per_group;
    const int tm = first_m + in_group % gsize;
    const int tn = in_group / gsize;
    const int m0 = tm * BM;
    const int n0 = tn * BN;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wm = wave >> 1;
    const int wn = wave & 1;

    T ra[LOADS], rb[LOADS];
    auto load_tile = [&](int k0) {
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            const int e = tid + THREADS * i;
            const int r = e >> 4;
            const int c = e & 15;
            const int gk = k0 + c;
            const int gm = m0 + r;
            const int gn = n0 + r;
            ra[i] = (gm < M && gk < K) ? A[(long long)gm * lda + gk] : T(0);
            rb[i] = (gn < N && gk < K) ? B[(long long)gn * ldb + gk] : T(0);
        }
    };
    auto store_tile = [&](int buf) {
#pragma unroll
        for (int i = 0; i < LOADS; ++i) {
            const int e = tid + THREADS * i;
            const int r = e >> 4;
            const int c = e & 15;
            As[buf][r * PITCH + c] = ra[i];
            Bs[buf][r * PITCH + c] = rb[i];
        }
    };

    typename MF::acc acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MF::zero();

    const int nk = afq::ceil_div(K, BK);
    load_tile(0);
    store_tile(0);
    __syncthreads();

    const int lr = lane & 15;
    const int lk = (lane >> 4) * 4;
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load_tile((kt + 1) * BK);

#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            T a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[buf][(wm * 64 + i * 16 + lr) * PITCH + lk + kk];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[buf][(wn * 64 + j * 16 + lr) * PITCH + lk + kk];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = MF::fma(a[i], b[j], acc[i][j]);
        }

        if (kt + 1 < nk) store_tile(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int col = n0 + wn * 64 + j * 16 + lr;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wm * 64 + i * 16 + MF::row(lane, r);
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

}  // namespace

namespace afq {

template <typename T, typename TOut>
int gemm_nt(const char* name, int M, int N, int K, T alpha, const T* A, long long lda, const T* B,
            long long ldb, T beta, TOut* C, long long ldc, hipStream_t stream) {
    AFQ_REQUIRE(M >= 0 && N >= 0 && K >= 0, "%s: negative size (M=%d N=%d K=%d)", name, M, N, K);
    if (M == 0 || N == 0) return 0;
    AFQ_REQUIRE(A && B && C, "%s: null pointer", name);
    AFQ_REQUIRE(lda >= K && ldb >= K && ldc >= N, "%s: leading dimension too small", name);
    const int tiles_m = ceil_div(M, BM);
    const int tiles_n = ceil_div(N, BN);
    const long long total = (long long)tiles_m * tiles_n;
    AFQ_REQUIRE(total < (1LL << 31), "%s: problem too large", name);
    hipLaunchKernelGGL((gemm_nt_kernel<T, TOut>), dim3((unsigned)total), dim3(THREADS), 0, stream, M, N, K,
                       alpha, A, lda, B, ldb, beta, C, ldc, tiles_m, tiles_n);
    return check_launch(name);
}

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
