// Shared definitions for the gfx950 phaseless-AFQMC kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace afq {

// Thread-local last-error text, read back through afq_last_error().
void set_error(const char* fmt, ...);
// Raises the dynamic-LDS limit of kernel `fn` to the 160 KB of a gfx950 CU,
// once per kernel and process (so no such call happens inside a stream capture
// after the first, eager launch).
void allow_full_lds(const void* fn);
// Converts a pending launch error into a return code (0 ok, -2 launch failure).
int check_launch(const char* what);
// Compute units of the current device (cached per device; 256 if the query fails).
int cu_count();

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ inline int round_up(int a, int b) { return ceil_div(a, b) * b; }

// Bijective remap of a linear workgroup id so that each XCD receives a
// contiguous range of tiles (blocks b, b+8, b+16, ... run on one XCD).
__device__ inline int xcd_contiguous(int b, int total) {
    const int x = b % kXcds;
    const int local = b / kXcds;
    const int base = total / kXcds;
    const int rem = total % kXcds;
    const int offset = x * base + (x < rem ? x : rem);
    return offset + local;
}

// Column tile of position nn in the order 0, nt - 1, 1, nt - 2, ... (a
// bijection of [0, nt)): with a B whose rows are zero past the diagonal, tile
// j's k-loop runs to its last column, and any contiguous range of positions
// (one XCD's share) then holds as much work as any other.
__device__ inline int zigzag(int nn, int nt) { return (nn & 1) ? nt - 1 - (nn >> 1) : nn >> 1; }

// v_mfma_f64_16x16x4_f64.  Lane l supplies A[l%16][l/16] and B[l/16][l%16];
// the accumulator holds D[4*r + l/16][l%16] in element r (measured on gfx950).
__device__ __forceinline__ f64x4 mfma_f64(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// v_mfma_f32_16x16x4_f32: same operand layout as the fp64 variant, but the
// accumulator holds D[4*(l/16) + r][l%16] in element r (as do the bf16 MFMAs).
__device__ __forceinline__ f32x4 mfma_f32(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Scalar-type dispatch for kernels templated on fp64 / fp32 arithmetic.
template <typename T>
struct mfma16;
template <>
struct mfma16<double> {
    using acc = f64x4;
    __device__ static __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
    __device__ static __forceinline__ acc fma(double a, double b, acc c) { return mfma_f64(a, b, c); }
    // row within the 16x16 tile of accumulator element r of lane `lane`
    __device__ static __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
};
template <>
struct mfma16<float> {
    using acc = f32x4;
    __device__ static __forceinline__ acc zero() { return acc{0.0f, 0.0f, 0.0f, 0.0f}; }
    __device__ static __forceinline__ acc fma(float a, float b, acc c) { return mfma_f32(a, b, c); }
    __device__ static __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// Four consecutive elements (16 or 32 bytes, naturally aligned).
template <typename T>
struct vec4;
template <>
struct vec4<double> {
    using type = double4;
};
template <>
struct vec4<float> {
    using type = float4;
};

// Index of the last group whose first item start(i) is <= x, for ascending
// starts with start(0) == 0 (binary search; x is uniform across the
// workgroup, so the loads are scalar).
template <typename F>
__device__ __forceinline__ int find_group(int count, int x, F start) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start(mid) <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// bf16 bits of x, round to nearest even (NaN -> quiet NaN): one hardware
// conversion (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

__device__ __forceinline__ double bf16_to_f64(uint16_t h) { return (double)__uint_as_float(((uint32_t)h) << 16); }

// The split-bf16 pair of an fp64 value: hi = bf16(x), lo = bf16(x - hi), both
// rounded to nearest even through fp32.
__device__ __forceinline__ void split_bf16(double x, uint16_t& hi, uint16_t& lo) {
    hi = f32_to_bf16_rne((float)x);
    lo = f32_to_bf16_rne((float)(x - bf16_to_f64(hi)));
}

// An fp32 value stored as its split-bf16 pair in 4 bytes (the size of the
// float it replaces): hi = bf16(v) in the low half, lo = bf16(v - hi) in the
// high half, both rounded to nearest even.  The split a bf16x3 consumer would
// otherwise compute on load, done once by the producer.
struct bf16_pair {
    uint32_t bits;
    bf16_pair() = default;
    __device__ explicit bf16_pair(float v) {
        const __bf16 h = (__bf16)v;
        const __bf16 l = (__bf16)(v - (float)h);
        bits = (uint32_t)__builtin_bit_cast(uint16_t, h) | ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
    }
};

// ---------------------------------------------------------------------------
// Symmetric VHS epilogue.  The VHS GEMM can run on the upper 16 x 16 blocks of
// the n x n plane only (V = sum_g x_g L_g is symmetric): its column
// c = blk * 256 + pi * 16 + qi is entry (bi * 16 + pi, bj * 16 + qi) of block
// blk = (bi, bj), bi <= bj, numbered row by row.  The GEMM then stores every
// entry twice into the full row-major (n, ldv) matrix of its output row: at
// (p, q) and, off the diagonal blocks, at (q, p).
struct sym_dst {
    int n, ldv, nb;  // nb = ceil(n / 16)
};

// (bi, bj) of upper block blk (uniform per tile: a short scalar loop)
__device__ __forceinline__ void sym_block(int blk, int nb, int& bi, int& bj) {
    int b = blk, i = 0;
    while (b >= nb - i) {
        b -= nb - i;
        ++i;
    }
    bi = i;
    bj = i + b;
}

// One lane's values v[j] = entry (bi*16 + pi0 + j, bj*16 + qi), j < NPI, of one
// output row (row points at that row's n x ldv matrix).  The mirror entries
// (bj*16 + qi, bi*16 + pi0 + j) are consecutive and go out in vector stores of
// W = min(NPI, 16 / sizeof(T)) elements: pi0 a multiple of W, ldv a multiple of
// 4 and the row base 16-byte aligned.  Entries past n in the stored
// blocks are zero (zero operand rows), so the padding columns n..ldv-1 get
// zeros from the direct stores; mirror rows past n are skipped.
template <int NPI, typename T>
__device__ __forceinline__ void sym_store(T* __restrict__ row, const sym_dst& d, int bi, int bj, int pi0, int qi,
                                          const T (&v)[NPI]) {
    const int q = bj * 16 + qi;
    if (q < d.ldv) {
#pragma unroll
        for (int j = 0; j < NPI; ++j) {
            const int p = bi * 16 + pi0 + j;
            if (p < d.n) row[(long long)p * d.ldv + q] = v[j];
        }
    }
    if (bi != bj && q < d.n) {
        constexpr int W = (NPI * (int)sizeof(T) >= 16) ? 16 / (int)sizeof(T) : NPI;
        typedef T vec_t __attribute__((ext_vector_type(W)));
        vec_t* dst = reinterpret_cast<vec_t*>(row + (long long)q * d.ldv + bi * 16 + pi0);
#pragma unroll
        for (int c = 0; c < NPI / W; ++c) {
            vec_t x;
#pragma unroll
            for (int e = 0; e < W; ++e) x[e] = v[c * W + e];
            dst[c] = x;
        }
    }
}

// symmetric-VHS GEMM in fp64 / fp32 (defined with the GEMMs)
template <typename T>
int gemm_nt_sym(const char* name, int M, int K, const T* A, long long lda, const T* B, long long ldb,
                sym_dst d, T* C, hipStream_t stream);

}  // namespace afq

#define AFQ_REQUIRE(cond, ...)         \
    do {                               \
        if (!(cond)) {                 \
            afq::set_error(__VA_ARGS__); \
            return -1;                 \
        }                              \
    } while (0)
