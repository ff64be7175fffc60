/*
 * C ABI of the gfx950 phaseless-AFQMC walker step (libafqmc_hip.so).
 *
 * Conventions
 *  - every pointer argument is a device pointer owned by the caller
 *    (except where noted); nothing is allocated inside the library;
 *  - row-major storage; complex numbers are complex128 interleaved (re, im)
 *    where the reference holds complex arrays, or split real planes
 *    ("planes (2, ...)": all real parts, then all imaginary parts) for the
 *    internal GEMM operands;
 *  - `stream` is a hipStream_t; all work is enqueued on it, nothing synchronizes;
 *  - return 0 on success, -1 on an invalid argument, -2 on a launch failure;
 *    afq_last_error() gives the message (thread-local).
 *
 * Reference paths are relative to the reference package ad_afqmc_prototype/.
 * Sizes: nw walkers, n orbitals, o occupied orbitals per spin, naux Cholesky
 * vectors, nspin = 1 (restricted walkers) or 2 (unrestricted walkers).
 */
#ifndef AFQMC_HIP_H
#define AFQMC_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message raised on the calling thread ("" if none). */
const char* afq_last_error(void);

/* ABI version of this header (bumped on any signature change). */
int afq_abi_version(void);

/*
 * fp64 GEMM  C = alpha * A * B^T + beta * C  on v_mfma_f64_16x16x4.
 * A (M x K, lda), B (N x K, ldb), C (M x N, ldc).  Building block of the force
 * bias and VHS contractions below; the reference has no standalone GEMM (XLA
 * lowers jnp.matmul, meas/rhf.py:33 and prop/chol_afqmc_ops.py:72-73).
 */
int afq_dgemm_nt(int M, int N, int K, double alpha, const double* A, long long lda,
                 const double* B, long long ldb, double beta, double* C, long long ldc,
                 void* stream);

/*
 * fp64 GEMM  C = alpha * A * B^T  with K cut into `splits` chunks (1..16, each
 * a multiple of 16 deep) run side by side: each chunk's partial product goes
 * to workspace (splits * M * N doubles, caller-owned), then the chunks are
 * summed in order (deterministic).  For problems with few output tiles and a
 * long K, the force bias of the fp64 step (meas/rhf.py:27-33, 60-72: A = the
 * Green's-function rows, B = the rotated Cholesky rows, K = nspin*o*n).
 */
int afq_dgemm_nt_splitk(int M, int N, int K, double alpha, const double* A, long long lda,
                        const double* B, long long ldb, double* C, long long ldc, int splits,
                        double* workspace, void* stream);

/*
 * fp32 GEMM  C = alpha * A * B^T + beta * C  on v_mfma_f32_16x16x4_f32, same
 * layout as afq_dgemm_nt; C is float32, or float64 when c_fp64 != 0.
 * Force-bias / VHS contractions of the mixed step in fp32 mode (meas/rhf.py:33,
 * prop/chol_afqmc_ops.py:72-73 in fp32).
 */
int afq_sgemm_nt(int M, int N, int K, float alpha, const float* A, long long lda,
                 const float* B, long long ldb, float beta, void* C, long long ldc,
                 int c_fp64, void* stream);

/*
 * Split-bf16 GEMM  C = alpha * (Ahi Bhi^T + Ahi Blo^T + Alo Bhi^T)  with fp32
 * accumulation (3 x v_mfma_f32_16x16x32_bf16 per step).  A (M x K) and B
 * (N x K) are given as bf16 hi/lo planes (uint16 bit patterns) from
 * afq_split_bf16; K, lda, ldb multiples of 8.  C is float32, or float64 when
 * c_fp64 == 1, or split-bf16 pair words (bf16(v) | bf16(v - bf16(v)) << 16, v
 * the fp32 result) when c_fp64 == 2.  Force-bias / VHS contractions of the mixed step in bf16x3
 * mode (meas/rhf.py:33, prop/chol_afqmc_ops.py:72-73).
 */
int afq_gemm_bf16x3_nt(int M, int N, int K, float alpha, const unsigned short* A_hi,
                       const unsigned short* A_lo, long long lda, const unsigned short* B_hi,
                       const unsigned short* B_lo, long long ldb, void* C, long long ldc,
                       int c_fp64, void* stream);

/*
 * The GEMMs above for a B whose row j is zero past column j (B[j][k] = 0 for
 * k > j; those entries must hold zeros, the tile on the diagonal still reads
 * them): every tile of output columns n0 .. n0 + bn - 1 stops its k-loop at
 * k = n0 + bn, which skips about half the products when N = K.  beta = 0.
 * Same results as the full-K call on the same B, bit for bit (the skipped
 * products are zeros).  The exchange's quadratic form vec(G)^T Y vec(G)
 * (meas/rhf.py:36-50, 75-105 regrouped) with Y symmetric equals
 * 2 vec(G)^T L^T vec(G) for L the lower triangle of Y with its diagonal halved:
 * Z = 2 G L^T on these entry points, half the products of Z = G Y.
 */
int afq_dgemm_nt_lower(int M, int N, int K, double alpha, const double* A, long long lda, const double* B,
                       long long ldb, double* C, long long ldc, void* stream);
int afq_sgemm_nt_lower(int M, int N, int K, float alpha, const float* A, long long lda, const float* B,
                       long long ldb, void* C, long long ldc, int c_fp64, void* stream);
int afq_gemm_bf16x3_nt_lower(int M, int N, int K, float alpha, const unsigned short* A_hi,
                             const unsigned short* A_lo, long long lda, const unsigned short* B_hi,
                             const unsigned short* B_lo, long long ldb, void* C, long long ldc, int c_fp64,
                             void* stream);

/*
 * afq_gemm_bf16x3_nt with the K range cut into `splits` chunks (whole 32-wide
 * k-steps) that run side by side: for problems of few output tiles, which then
 * fill the GPU.  Each chunk writes its fp32 partial product to workspace
 * (splits x M x N floats, caller-owned), and a second kernel forms
 * C = alpha * (sum of the partials in chunk order, in fp64): deterministic.
 * Same operand rules as afq_gemm_bf16x3_nt; 1 <= splits <= 16.
 */
int afq_gemm_bf16x3_nt_splitk(int M, int N, int K, float alpha, const unsigned short* A_hi,
                              const unsigned short* A_lo, long long lda, const unsigned short* B_hi,
                              const unsigned short* B_lo, long long ldb, void* C, long long ldc, int c_fp64,
                              int splits, float* workspace, void* stream);

/*
 * x (rows x cols, float64, row pitch ld_src) -> hi = bf16(x), lo = bf16(x - hi)
 * (rows x ld_dst, zero in columns >= cols).  Operand preparation for
 * afq_gemm_bf16x3_nt; no counterpart in the reference (it computes in fp64).
 */
int afq_split_bf16(int rows, int cols, const double* src, long long ld_src,
                   unsigned short* hi, unsigned short* lo, long long ld_dst, void* stream);

/*
 * float64 (rows x cols, pitch ld_src) -> float32 (rows x ld_dst, zero padded).
 * Operand preparation for the fp32 kernels; no counterpart in the reference.
 */
int afq_convert_f32(int rows, int cols, const double* src, long long ld_src, float* dst,
                    long long ld_dst, void* stream);

/*
 * Overlap matrix m = C^T w, LU with partial pivoting, log det(m), and the
 * half-rotated Green's function G = (w m^-1)^T, per walker and spin.
 * Replaces _half_green_from_overlap_matrix (meas/rhf.py:15-19), overlap_r
 * (meas/rhf.py:22-24, trial/rhf.py:35-37), overlap_u (meas/rhf.py:53-57,
 * trial/rhf.py:40-44) and _det (trial/rhf.py:25-26), with the determinant
 * carried as a complex logarithm.
 *   walkers_s : complex128 (nw, n, o)      (walkers_dn NULL when nspin == 1)
 *   trial_s   : float64 (n, o), real trial orbitals
 *   greens    : planes (2, nw, nspin*o*n); element (s, i, p) at (s*o + i)*n + p
 *   logdet    : complex128 (nw, nspin), log det(C_s^T w_s)
 */
int afq_overlap_greens(int nw, int n, int o, int nspin,
                       const double* walkers_up, const double* walkers_dn,
                       const double* trial_up, const double* trial_dn,
                       double* greens, double* logdet, void* stream);

/*
 * afq_overlap_greens with complex trial orbitals (trial/rhf.py:25-44 takes
 * C^H): trial_up / trial_dn are complex128 (n, o) interleaved (re, im), and
 * m = C^H w.  fp64 rows and log det as afq_overlap_greens; o <= 64.
 */
int afq_overlap_greens_ctrial(int nw, int n, int o, int nspin, const double* walkers_up,
                              const double* walkers_dn, const double* trial_up, const double* trial_dn,
                              double* greens, double* logdet, void* stream);

/*
 * The bf16x3 step's Green's function (meas/rhf.py:15-33, trial/rhf.py:25-37):
 * log det(C_s^T w_s) in fp64 as afq_overlap_greens (partial pivoting on
 * |re| + |im|, the LU's pivots, here in a Gauss-Jordan inversion), and the
 * rows of G = (w m^-1)^T written only as their split-bf16 planes g_hi / g_lo
 * (afq_overlap_greens_split's layout: 2 nw rows of ldk, ldk in
 * [nspin*o*n, nspin*o*n + 64), padding zeroed), formed on the bf16 matrix
 * cores as (split m^-1)^T (split w)^T: accurate to the split-bf16 product
 * (~2^-16 relative), the precision the bf16x3 force bias reads them at.  No
 * fp64 rows.  o <= 32.
 */
int afq_overlap_greens_bf16x3(int nw, int n, int o, int nspin, const double* walkers_up,
                              const double* walkers_dn, const double* trial_up, const double* trial_dn,
                              double* logdet, unsigned short* g_hi, unsigned short* g_lo, long long ldk,
                              void* stream);

/*
 * afq_overlap_greens that also writes the Green's function rows as split-bf16
 * planes g_hi / g_lo (2 nw rows of ldk: real rows, then imaginary rows, the
 * values afq_split_bf16 makes of greens; columns nspin o n .. ldk - 1 zeroed),
 * the A operand of the bf16x3 force-bias contraction, without a second pass
 * over greens.  nspin o n <= ldk < nspin o n + 64.
 */
int afq_overlap_greens_split(int nw, int n, int o, int nspin, const double* walkers_up, const double* walkers_dn,
                             const double* trial_up, const double* trial_dn, double* greens, double* logdet,
                             unsigned short* g_hi, unsigned short* g_lo, long long ldk, void* stream);

/*
 * Force bias fb = alpha * rot . vec(G): planes (2, nw, naux).
 * rot (naux, k) holds the half-rotated Cholesky vectors of every spin sector
 * side by side (k = nspin*o*n).  alpha = 2 for restricted walkers.
 * Replaces force_bias_kernel_r / _u (meas/rhf.py:27-33, 60-72).
 */
int afq_force_bias(int nw, int naux, int k, double alpha, const double* greens,
                   const double* rot, double* fb, void* stream);

/*
 * Field shifts, shifted fields and the two per-walker sums of
 * prop/afqmc.py:97-101:
 *   fs = -sqrt_dt (i fb - mf);  xbar = x - fs;
 *   shift_term = sum_g xbar_g mf_g;  fb_term = sum_g (x_g fs_g - fs_g^2 / 2).
 *   fields (nw, naux) float64; fb planes (2, nw, naux); mf complex128 (naux);
 *   xbar planes (2, nw, naux); shift_term, fb_term complex128 (nw).
 */
int afq_field_shifts(int nw, int naux, double sqrt_dt, const double* fields,
                     const double* fb, const double* mf, double* xbar,
                     double* shift_term, double* fb_term, void* stream);

/*
 * afq_field_shifts that also writes the shifted fields as split-bf16 planes
 * x_hi / x_lo (2 nw rows of ldk: real rows, then imaginary rows, the values
 * afq_split_bf16 makes of xbar; columns naux .. ldk - 1 zeroed): the A operand
 * of the bf16x3 VHS contraction, without a second pass over xbar.
 */
int afq_field_shifts_split(int nw, int naux, double sqrt_dt, const double* fields, const double* fb,
                           const double* mf, double* xbar, double* shift_term, double* fb_term,
                           unsigned short* x_hi, unsigned short* x_lo, long long ldk, void* stream);

/*
 * VHS = sum_g xbar_g L_g, planes (2, nw, n, ldv) (ldv = n rounded up to 4).
 * chol_t (n*ldv, naux) is L transposed with zero rows at q >= n.
 * Replaces _make_vhs_split_flat (prop/chol_afqmc_ops.py:70-74).
 */
int afq_vhs(int nw, int n, int ldv, int naux, const double* xbar, const double* chol_t,
            double* vhs, void* stream);

/*
 * Symmetric VHS contraction, the VHS GEMM of _make_vhs_split_flat
 * (prop/chol_afqmc_ops.py:70-74) on half the work: V = sum_g x_g L_g is
 * symmetric, so A (rows x k: shifted-field rows) is contracted only against
 * the Cholesky rows of the upper 16 x 16 blocks of the (p, q) plane, B
 * (nb (nb + 1) / 2 * 256 rows x k, nb = ceil(n / 16)): block (bi, bj),
 * bi <= bj, numbered row by row, element (pi, qi) at row 256 blk + 16 pi + qi,
 * entry (16 bi + pi, 16 bj + qi), zero rows past n.  Each result goes straight
 * to both of its places in vhs (rows, n, ldv), 16-byte aligned; columns
 * n..ldv-1 receive zeros.
 *   arith 0: fp64 (a_hi, b_hi double; vhs double)
 *   arith 1: fp32 (a_hi, b_hi float; vhs float)
 *   arith 2: bf16x3 (hi / lo planes of afq_split_bf16; k, lda, ldb multiples
 *            of 8; vhs float)
 */
int afq_vhs_sym(int rows, int n, int ldv, int k, int arith, const void* a_hi, const void* a_lo,
                long long lda, const void* b_hi, const void* b_lo, long long ldb, void* vhs, void* stream);

/*
 * Trotter step w' = E . sum_{k<=n_terms} (i sqrt_dt VHS)^k / k! . E w per walker
 * and spin (both spins share the walker's VHS), E = exp(-dt/2 h1_eff) (n, ldv).
 * Replaces _apply_trotter_r / _u (prop/chol_afqmc_ops.py:106-116, 127-154,
 * 173-196) and taylor_expm_action (prop/utils.py:20-31).
 *   win_s, wout_s : complex128 (nw, n, o); wout must not alias win.
 */
int afq_trotter(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                const double* exp_h1_half, const double* vhs,
                const double* win_up, const double* win_dn,
                double* wout_up, double* wout_dn, void* stream);

/*
 * Adjoint of the Green's function and log-overlap (meas/rhf.py:15-24,
 * trial/rhf.py:25-44 differentiated; restricted walkers, real trial C (n, o),
 * o <= 32, n <= 128 with o = 32): for G = (w m^-1)^T (o x n), ld = log det m, m = C^T w, and the
 * complex adjoints g_bar ((nw, o, n) complex128, may be null) and ld_bar ((nw)
 * complex128, may be null; not both), writes (accumulate = 0) or adds
 * (accumulate = 1) w_bar = [g_bar^T - C (m^-H w^H g_bar^T - ld_bar I)] m^-H
 * into w_bar ((nw, n, o) complex128).  m^-1 by Gauss-Jordan elimination with
 * partial pivoting.  No counterpart in the reference (jax differentiates it).
 */
int afq_greens_adjoint(int nw, int n, int o, const double* walkers, const double* trial, const double* g_bar,
                       const double* ld_bar, double* w_bar, int accumulate, void* stream);

/*
 * Adjoint of afq_trotter (fp64, restricted walkers, n <= 128, o <= 32) for
 * the reverse pass of a differentiated run (prop/chol_afqmc_ops.py:173-196,
 * prop/utils.py:20-31 differentiated): w' = E u, u = sum_k t_k, t_0 = E w,
 * t_k = (i sqrt_dt / k) vhs t_{k-1}.  Given w_bar_out (dJ/dw' as complex
 * adjoints, (nw, n, o) complex128), writes w_bar_in = E^T t_bar_0, vhs_bar
 * (planes (2, nw, n, n): sum_k conj(i sqrt_dt / k) t_bar_k t_{k-1}^H) and
 * e_half_bar (nw, n, n): each walker's share Re(w_bar_out u^H + t_bar_0 w^H)
 * of E's adjoint (the caller sums them).  exp_h1_half and exp_h1_half_t (its
 * transpose) are (n, ldv); vhs the forward step's planes (2, nw, n, ldv)
 * (complex symmetric); workspace 2 (2 n_terms + 2) nw n 16 ceil(o / 16)
 * doubles.  No counterpart in the reference (jax differentiates its step).
 */
int afq_trotter_adjoint(int nw, int n, int o, int ldv, int n_terms, double sqrt_dt, const double* exp_h1_half,
                        const double* exp_h1_half_t, const double* vhs, const double* w_in, const double* w_bar_out,
                        double* w_bar_in, double* vhs_bar, double* e_half_bar, double* workspace, void* stream);

/*
 * afq_trotter with E, VHS and the Taylor recursion in fp32 (walkers still read
 * and written as complex128): the Trotter step of the mixed step.
 * exp_h1_half (n, ldv) float32; vhs planes (2, nw, n, ldv) float32.
 */
int afq_trotter_f32(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                    const float* exp_h1_half, const float* vhs,
                    const double* win_up, const double* win_dn,
                    double* wout_up, double* wout_dn, void* stream);

/*
 * afq_trotter_f32 with every product on split-bf16 MFMAs (x = hi + lo, three
 * bf16 products accumulated in fp32), E applied as X + (E - I) X with X added
 * exactly; running sums in fp32.  Same arguments as afq_trotter_f32, same
 * reference functions.  n > 224 runs the fp32 step instead.
 */
int afq_trotter_bf16x3(int nw, int n, int o, int nspin, int ldv, int n_terms, double sqrt_dt,
                       const float* exp_h1_half, const float* vhs,
                       const double* win_up, const double* win_dn,
                       double* wout_up, double* wout_dn, void* stream);

/*
 * afq_trotter_bf16x3 with V and E - I (E = exp(-dt h1 / 2); both symmetric)
 * given as their upper 16 x 16 blocks: nb (nb + 1) / 2 * 256 entries per
 * matrix, nb = ceil(n / 16), in afq_vhs_sym's column order (block (r, c),
 * r <= c, numbered row by row, entry (pi, qi) at 256 blk + 16 pi + qi),
 * entries past n zero.  Every entry is a split-bf16 pair word: bf16(v) in the
 * low 16 bits, bf16(v - bf16(v)) in the high 16 bits, v the fp32 value
 * (afq_gemm_bf16x3_nt with c_fp64 = 2 writes V so).  vhs_sym holds two planes
 * (real, imaginary) of nw such rows: the product of afq_gemm_bf16x3_nt with
 * the upper-block Cholesky rows, 0.58 of the bytes of full rows at n = 200.
 * n <= 224 (the on-chip capacity; larger n is an error).
 */
int afq_trotter_bf16x3_sym(int nw, int n, int o, int nspin, int n_terms, double sqrt_dt,
                           const unsigned* exp_h1_half_sym, const unsigned* vhs_sym, const double* win_up,
                           const double* win_dn, double* wout_up, double* wout_dn, void* stream);

/*
 * afq_trotter_bf16x3_sym with the two one-body halves given separately:
 *   W_out = F_last . Taylor_m(i sqrt_dt V) . F_first . W_in,
 * e_first = F_first - I and e_last = F_last - I in the same upper-block
 * pair-word layout; e_last NULL leaves out the last half (W_out = Taylor .
 * F_first . W_in).  Consecutive steps of prop/chol_afqmc_ops.py:173-196 apply
 * E . E between two Taylor series; with e_last NULL on one step and e_first =
 * E^2 - I (E^2 = exp(-dt h1)) on the next, the pair becomes one product.  A
 * walker left without its last half is E^-1 W: its overlap matrix is
 * (E C)^T (E^-1 W) with the trial E C, and its Green's function G E^-1, so
 * the force bias contracts it with C^T L_g E.  afq_trotter_bf16x3_sym(..., e,
 * ...) is afq_trotter_bf16x3_sym_halves(..., e, e, ...).
 */
int afq_trotter_bf16x3_sym_halves(int nw, int n, int o, int nspin, int n_terms, double sqrt_dt,
                                  const unsigned* e_first, const unsigned* e_last, const unsigned* vhs_sym,
                                  const double* win_up, const double* win_dn, double* wout_up, double* wout_dn,
                                  void* stream);

/*
 * Phaseless importance, cos-phase cut, floor/cap and population-control shift
 * of prop/afqmc.py:116-138, on log-overlaps:
 *   log ovlp_new = (nspin == 1 ? 2 : 1) * sum_s logdet_new[w, s]
 *   ratio = exp(log ovlp_new - log ovlp_old) ...
 *   pop_shift_new = e_estimate - damping * log(max(mean w_new, 1e-300)) / dt.
 * logdet_new complex128 (nw, nspin); logovlp_* complex128 (nw); shift_term,
 * fb_term complex128 (nw); weights_* float64 (nw); pop_shift_* device float64[1].
 */
int afq_phaseless_update(int nw, int nspin, const double* logdet_new,
                         const double* logovlp_old, const double* shift_term,
                         const double* fb_term, const double* weights_old,
                         const double* pop_shift_old, double e_estimate,
                         double h0_prop_re, double h0_prop_im, double dt, double sqrt_dt,
                         double weight_floor, double weight_cap, double damping,
                         double* logovlp_new, double* weights_new, double* pop_shift_new,
                         void* stream);

/*
 * Reduced QR with diag(R) real and positive, keeping Q (the norm is dropped),
 * per walker: complex128 (nw, n, o) -> (nw, n, o).
 * Replaces _qr / orthogonalize / orthonormalize (walkers.py:169-211).
 */
int afq_orthonormalize(int nw, int n, int o, const double* win, double* wout, void* stream);

/*
 * Local energy per walker, complex128 (nw):
 *   restricted:   h0 + 2 sum G.rot_h1 + 2 sum_g c_g^2 - sum_g tr(f_g f_g)
 *   unrestricted: h0 + sum_s G_s.rot_h1_s + [sum_g (c_up+c_dn)^2 - x_up - x_dn] / 2
 * with f_g = rot_chol_g G^T (o x o) and c_g = tr f_g, never materialised.
 *   rot_chol_pad (nspin, naux, opad, ldr) float64, opad = o rounded up to 16,
 *   ldr = n rounded up to 4, zero padded; rot_h1 (nspin, o, n) float64.
 * Replaces energy_kernel_r / _u (meas/rhf.py:36-50, 75-105).
 */
int afq_energy(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
               const double* rot_chol_pad, int ldr, const double* rot_h1, double* energy,
               void* stream);

/*
 * afq_energy with the f_g contractions on fp32 MFMA (rot_chol_pad float32) and
 * every sum over g (Coulomb squares, exchange, one-body) in fp64; two walkers
 * share each streamed rot_chol_g tile when they fit in LDS.  The measurement
 * of the fp32 mixed step.
 */
int afq_energy_f32(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
                   const float* rot_chol_pad, int ldr, const double* rot_h1, double* energy,
                   void* stream);

/*
 * afq_energy with the f_g contractions as split-bf16 products (rot_chol and G
 * both carried as bf16 hi + lo, three bf16 MFMAs, fp32 sums within a g) and
 * every sum over g in fp64.  rot_hi / rot_lo: bf16 bit planes (nspin, naux,
 * opad, ldr) with opad = o rounded up to 16, zero beyond (o, n); ldr a
 * multiple of 32, n <= ldr <= 256.  Same reference functions as afq_energy.
 */
int afq_energy_bf16x3(int nw, int n, int o, int nspin, int naux, double h0, const double* greens,
                      const unsigned short* rot_hi, const unsigned short* rot_lo, int ldr,
                      const double* rot_h1, double* energy, void* stream);

/*
 * Block energy of prop/blocks.py:59-79: real part, outlier clip to e_ref when
 * |e - e_ref| > thresh, e_block = sum w e / sum w (e_ref when sum w == 0), and
 * pop_shift_new = (1 - shift_ema) pop_shift + shift_ema e_block.
 *   out (device float64[3]) = {e_block, sum w, pop_shift_new}.
 */
int afq_block_energy(int nw, const double* energy, const double* weights, double e_ref,
                     double thresh, double shift_ema, const double* pop_shift, double* out,
                     void* stream);

/*
 * Comb reconfiguration indices of _sr_indices / stochastic_reconfiguration
 * (walkers.py:233-265): sequential left-to-right cumulative sum of |w|,
 * z_k = tot * (k + zeta) / nw, idx_k = first i with cw_i >= z_k,
 * weights_new = tot / nw.  cumsum (nw) is caller scratch; idx int32 (nw).
 */
int afq_comb_indices(int nw, const double* weights, double zeta, double* cumsum, int* idx,
                     double* weights_new, void* stream);

/*
 * afq_comb_indices with zeta read from device memory (zeta: one float64, e.g.
 * written by afq_rng_uniform on the same stream), so that the draw and the
 * comb need no host round trip.  Same reference functions.
 */
int afq_comb_indices_dev(int nw, const double* weights, const double* zeta, double* cumsum,
                         int* idx, double* weights_new, void* stream);

/*
 * afq_comb_indices_dev without the cumsum scratch argument: it allocates
 * nothing and takes no workspace.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The cumulative sum goes to a
 * static device array of the library (2^18 fp64 entries, 2 MiB per GPU, held
 * until the process ends); nw above 2^18 is rejected (-1).  Every call on one
 * GPU shares that array: calls that overlap on one GPU race.  Same results as
 * afq_comb_indices_dev.
 */
int afq_comb_indices_noalloc(int nw, const double* weights, const double* zeta, int* idx,
                             double* weights_new, void* stream);

/*
 * The step's random stream on the device (prop/afqmc.py:32,86-88,
 * prop/blocks.py:81-82: key, subkey = split(key); normal(subkey, shape) /
 * uniform(subkey)), in jax's original Threefry-2x32 scheme (config.py:27).
 * Keys are device uint32[2].
 *   afq_rng_split:   (new_key, subkey) = split(key); the three distinct.
 *   afq_rng_normal:  out[0..n) = sqrt(2) erfinv(uniform(nextafter(-1, 0), 1)),
 *                    value i from threefry(key, (i, n + i)); n < 2^31.
 *   afq_rng_uniform: out[0] = uniform in [0, 1) from threefry(key, (0, 1)).
 * The uniforms follow jax's bit recipe; erfinv is the library's own (Halley
 * iterations on erf / erfc), so normals match another erfinv to a few ulp.
 */
int afq_rng_split(const unsigned* key, unsigned* new_key, unsigned* subkey, void* stream);
int afq_rng_normal(const unsigned* key, long long n, double* out, void* stream);
int afq_rng_uniform(const unsigned* key, double* out, void* stream);

/*
 * The Threefry-2x32-20 block function itself (Salmon et al., SC 2011) on n
 * counter pairs (c0[i], c1[i]) under key (k0, k1): (o0[i], o1[i]).  No
 * reference counterpart (the reference reaches it through jax.random); exposed
 * for known-answer tests.
 */
int afq_threefry2x32(unsigned k0, unsigned k1, long long n, const unsigned* c0, const unsigned* c1,
                     unsigned* o0, unsigned* o1, void* stream);

/*
 * dst[p][k][:] = src[p][idx[k]][:] for p < nplanes, rows of row_len doubles.
 * The walker gather of stochastic_reconfiguration (walkers.py:258-263).
 */
int afq_gather_rows(int nrows, long long row_len, int nplanes, long long src_plane_stride,
                    long long dst_plane_stride, const int* idx, const double* src, double* dst,
                    void* stream);

/*
 * ---------------------------------------------------------------------------
 * Grouped launches: many independent systems (uneven n, o, naux and walker
 * counts, each with its own Hamiltonian, RHF trial and restricted walkers)
 * advance together, one launch per stage of the step for all of them.  The
 * reference runs one system per call; these entry points batch that call
 * (prop/afqmc.py:74-147 per system) and have no single counterpart there.
 *
 * The caller keeps a device array of descriptors, one per system, written
 * once: they hold the system's sizes, scalars and device pointers into the
 * caller's buffers.  Two tables (input and output buffers swapped) serve
 * alternate steps, so a sequence of steps needs no descriptor rewrite and can
 * be captured in a HIP graph.  The systems of one table must not share output
 * buffers.  `count` systems; `total_walkers` = sum of nw; every system has
 * n <= 256, o <= 32 and o <= n.
 * ---------------------------------------------------------------------------
 */
typedef struct afq_system_step {
    int n, o, naux, nw;          /* orbitals, occupied orbitals, Cholesky vectors, walkers */
    int ldv;                     /* n rounded up to 4: row pitch of exp_h1_half and VHS */
    int walker0;                 /* first walker of this system in the batch (prefix sum of nw) */
    int tblock0;                 /* first workgroup of the launch this table serves: prefix sum of
                                    nw * ceil(o / 16) (Trotter), of ceil(nw / wpb) (bf16x3 energy) */
    int ldr;                     /* row pitch of rot_chol: n rounded up to 4 (fp32), to 32 (bf16x3) */
    double sqrt_dt, dt, e_estimate, h0_prop;
    double weight_floor, weight_cap, damping, h0;
    const double* trial;         /* (n, o) float64 RHF orbitals */
    const double* walkers_in;    /* complex128 (nw, n, o) */
    double* walkers_out;         /* complex128 (nw, n, o) */
    double* greens;              /* planes (2, nw, o*n): G of walkers_out (Green's stage) */
    double* logdet;              /* complex128 (nw): log det(C^T walkers_out) */
    double* fields;              /* (nw, naux) float64 */
    const double* fb;            /* force bias planes (2, nw, naux) */
    const double* mf;            /* complex128 (naux) mean-field shifts */
    double* xbar;                /* shifted-field planes (2, nw, naux) */
    double* shift_term;          /* complex128 (nw) */
    double* fb_term;             /* complex128 (nw) */
    const float* exp_h1_half;    /* (n, ldv) float32; bf16x3 pair-word Trotter: E - I as upper-block pairs */
    const float* vhs;            /* planes (2, nw, n, ldv) float32; bf16x3 pair-word Trotter: (2, nw,
                                    nb (nb + 1) / 2 * 256) pair words of V's upper 16 x 16 blocks */
    const double* log_overlaps_in;  /* complex128 (nw) */
    double* log_overlaps_out;       /* complex128 (nw) */
    const double* weights_in;    /* (nw) */
    double* weights_out;         /* (nw) */
    const double* pop_in;        /* float64[1] population-control shift */
    double* pop_out;             /* float64[1] */
    const unsigned* key_in;      /* uint32[2] Threefry key */
    unsigned* key_out;           /* uint32[2] */
    const float* rot_chol;       /* (naux, round16(o), ldr) float32 half-rotated Cholesky, zero padded;
                                    bf16x3 energy: its hi bf16 plane, then the lo plane, same shape */
    const double* rot_h1;        /* (o, n) float64 */
    double* energy;              /* complex128 (nw) local energies of walkers_out */
    unsigned short* greens_hi;   /* bf16x3 step: (2, nw, ldk) bf16 hi plane of the Green's-function rows */
    unsigned short* greens_lo;   /* (2, nw, ldk) lo plane */
    long long ldk;               /* row pitch of the planes: o*n <= ldk < o*n + 64 */
    unsigned short* xbar_hi;     /* null, or (2, nw, ldx) bf16 hi plane of xbar written by the field shifts */
    unsigned short* xbar_lo;     /* (2, nw, ldx) lo plane */
    long long ldx;               /* row pitch of the xbar planes: >= naux */
    double* exch_z;              /* planes (2, nw, o*n): Z = G Y of the exchange as a quadratic form
                                    (afq_energy_assemble_grouped) */
    const float* exp_h1;         /* bf16x3 pair-word Trotter: E^2 - I (E^2 = exp(-dt h1_eff)) as upper-block
                                    pairs, the first half after a step without its last
                                    (afq_trotter_bf16x3_sym_grouped_halves) */
} afq_system_step;

/* One problem of a grouped GEMM: C (M x N) = alpha A (M x K) B^T (N x K). */
typedef struct afq_gemm_problem {
    int M, N, K;
    int tile_start;              /* first workgroup of this problem: prefix sum of its tiles */
    long long lda, ldb, ldc;
    const float* A;
    const float* B;
    void* C;                     /* float32, or float64 when c_fp64 != 0 */
    float alpha;
    int c_fp64;
    /* sym_n > 0: the symmetric VHS form of afq_vhs_sym: the N columns are the
     * upper 16 x 16 block entries of an sym_n x sym_n matrix (N = nb (nb + 1) / 2
     * * 256, nb = ceil(sym_n / 16)) and every result goes to both of its places
     * in C's rows of sym_n x sym_ldv floats (sym_ldv a multiple of 4, C float32
     * and 16-byte aligned; ldc unused) */
    int sym_n, sym_ldv;
    int b_lower;                 /* 1: row j of B is zero past column j, and every column tile's k-loop stops
                                    at its last column (as afq_sgemm_nt_lower) */
} afq_gemm_problem;

/*
 * Tile shape `shape` (0..3) of the grouped fp32 GEMM: *bm x *bn rows x
 * columns of C per workgroup.  A problem has ceil(M/bm) * ceil(N/bn) tiles.
 * Host-only query, no device work.
 */
int afq_gemm_tile_shape(int shape, int* bm, int* bn);

/*
 * Every problem's fp32 GEMM in one launch of total_tiles workgroups (the sum
 * of the problems' tiles for `shape`, tile_start their prefix sums).  The
 * force-bias and VHS contractions of every system of a grouped step
 * (meas/rhf.py:33, prop/chol_afqmc_ops.py:72-73 per system).
 */
int afq_sgemm_nt_grouped(int count, const afq_gemm_problem* problems, int total_tiles, int shape,
                         void* stream);

/* One problem of a grouped bf16x3 GEMM: C (M x N) = alpha A (M x K) B^T (N x K)
 * with A and B as bf16 hi / lo planes (afq_split_bf16); K, lda and ldb
 * multiples of 8, planes 16-byte aligned. */
typedef struct afq_gemm_problem_bf16x3 {
    int M, N, K;
    int tile_start;              /* first workgroup: prefix sum of the problems' tiles of the launch's shape */
    long long lda, ldb, ldc;
    const unsigned short* A_hi;
    const unsigned short* A_lo;
    const unsigned short* B_hi;
    const unsigned short* B_lo;
    void* C;                     /* c_fp64 0: float32, 1: float64, 2: split-bf16 pair words (afq_gemm_bf16x3_nt) */
    float alpha;
    int c_fp64;
    int sym_n, sym_ldv;          /* sym_n > 0: the symmetric VHS store of afq_gemm_problem */
    int b_lower;                 /* as afq_gemm_problem's (afq_gemm_bf16x3_nt_lower) */
} afq_gemm_problem_bf16x3;

/*
 * Every problem's bf16x3 GEMM in one launch of total_tiles workgroups, all in
 * tile shape `shape` (afq_gemm_bf16x3_tile_shape: 0 = 128 x 128 staged
 * through registers, 1 = 128 x 256 and 2 = 256 x 256 moved to LDS by
 * buffer_load ... lds; every problem's A and B planes below 2 GiB).  Each
 * element's sum runs in the same order in every shape, so the shapes give the
 * same bits.  The force-bias and VHS contractions of every system of a grouped
 * step in split-bf16 precision (meas/rhf.py:33, prop/chol_afqmc_ops.py:72-73
 * per system).
 */
int afq_gemm_bf16x3_nt_grouped(int count, const afq_gemm_problem_bf16x3* problems, int total_tiles, int shape,
                               void* stream);

/* Rows (*bm) and columns (*bn) of C per workgroup of afq_gemm_bf16x3_nt_grouped
 * tile shape `shape` (0..2); no counterpart in the reference (tile choice of
 * the batched step's autotuner). */
int afq_gemm_bf16x3_tile_shape(int shape, int* bm, int* bn);

/* One problem of afq_split_bf16_grouped: afq_split_bf16's arguments. */
typedef struct afq_split_problem {
    const double* src;
    unsigned short* hi;
    unsigned short* lo;
    int rows, cols;
    int ld_src, ld_dst;
} afq_split_problem;

/*
 * afq_split_bf16 of every problem in one launch (count <= 65535; max_elems =
 * the largest rows * ld_dst).  Operand preparation of the grouped bf16x3
 * GEMM; no counterpart in the reference.
 */
int afq_split_bf16_grouped(int count, const afq_split_problem* problems, long long max_elems, void* stream);

/*
 * Per system: key_out, sub = split(key_in); fields = normal(sub, (nw, naux))
 * (prop/afqmc.py:86-88), one launch for all systems; nw * naux < 2^31 each.
 */
int afq_rng_fields_grouped(int count, const afq_system_step* sys, int total_walkers, void* stream);

/* afq_field_shifts for every system (prop/afqmc.py:97-101); a system whose
 * xbar_hi is set also gets xbar's split-bf16 planes (afq_field_shifts_split). */
int afq_field_shifts_grouped(int count, const afq_system_step* sys, int total_walkers, void* stream);

/*
 * afq_trotter_f32 for every system: walkers_in -> walkers_out with the
 * system's exp_h1_half and VHS (prop/chol_afqmc_ops.py:173-196,
 * prop/utils.py:20-31).  total_blocks = sum of nw * ceil(o / 16); max_n the
 * largest n.
 */
int afq_trotter_f32_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n,
                            int n_terms, void* stream);

/*
 * afq_trotter_bf16x3 for every system of one bucket of equal ceil(round_up(n,
 * 16) / 32) (the workgroup size): walkers_in -> walkers_out with the system's
 * fp32 exp_h1_half and VHS in (n, ldv) rows (prop/chol_afqmc_ops.py:173-196,
 * prop/utils.py:20-31).  total_blocks = sum of nw * ceil(o / 16); max_n the
 * largest n of the bucket (<= 224).
 */
int afq_trotter_bf16x3_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n,
                               int n_terms, void* stream);

/*
 * afq_trotter_bf16x3_sym for every system of one bucket, like
 * afq_trotter_bf16x3_grouped: the descriptor's vhs holds the system's V as
 * split-bf16 pair words of its upper 16 x 16 blocks (2 planes of nw *
 * nb (nb + 1) / 2 * 256 words, nb = ceil(n / 16); the pair output of the VHS
 * GEMM) and exp_h1_half holds E - I in the same layout (prop/chol_afqmc_ops.py:173-196,
 * prop/utils.py:20-31).
 */
int afq_trotter_bf16x3_sym_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n,
                                   int n_terms, void* stream);

/*
 * afq_trotter_bf16x3_sym_grouped with the halves of afq_trotter_bf16x3_sym_halves
 * for every system: first_full != 0 takes each system's exp_h1 (E^2 - I) for the
 * first half instead of exp_h1_half, last_half == 0 leaves out the last half.
 */
int afq_trotter_bf16x3_sym_grouped_halves(int count, const afq_system_step* sys, int total_blocks, int max_n,
                                          int n_terms, int first_full, int last_half, void* stream);

/*
 * afq_overlap_greens for every system, of walkers_out into greens / logdet
 * (meas/rhf.py:15-24, trial/rhf.py:25-37).  max_o the largest o.
 */
int afq_overlap_greens_grouped(int count, const afq_system_step* sys, int total_walkers, int max_o,
                               void* stream);

/*
 * afq_overlap_greens_bf16x3 (restricted walkers, o <= 32) for every system:
 * log det(C^T walkers_out) into logdet and the Green's-function rows as the
 * force bias's split-bf16 planes into greens_hi / greens_lo (padding columns
 * zeroed); no fp64 rows (meas/rhf.py:15-24, trial/rhf.py:25-44 per system).
 * max_o the largest o of the table (the kernel's capacity: 16 or 32).
 */
int afq_overlap_greens_bf16x3_grouped(int count, const afq_system_step* sys, int total_walkers, int max_o,
                                      void* stream);

/* afq_phaseless_update for every system, one workgroup each (prop/afqmc.py:116-138). */
int afq_phaseless_update_grouped(int count, const afq_system_step* sys, void* stream);

/*
 * The block end of every system up to the comb, one workgroup each
 * (prop/blocks.py:59-92; per system afq_block_energy, afq_rng_split,
 * afq_rng_uniform and afq_comb_indices_dev, with the same results bit for
 * bit): from energy (real parts), weights_in, pop_in, e_estimate and key_in,
 * stats[2 i] = the clipped block energy of system i (local energies farther
 * than thresh from e_estimate count as e_estimate), stats[2 i + 1] = its weight
 * sum; in place pop_in = (1 - shift_ema) pop_in + shift_ema stats[2 i],
 * (key_in, sub) = split(key_in), weights_in = the comb's weights for
 * zeta = uniform(sub).  idx (int32) and cw (float64, caller scratch): one
 * entry per walker of the batch at walker0; idx gets the comb indices.
 */
int afq_block_end_grouped(int count, const afq_system_step* sys, double thresh, double shift_ema, double* stats,
                          double* cw, int* idx, void* stream);

/*
 * The reconfiguration gather of every system, one workgroup per walker
 * (prop/blocks.py:81-92, walkers.py:233-265 per system; afq_gather_rows per
 * array): with j = idx[walker0 + k], walkers_in[k] = walkers_out[j],
 * log_overlaps_in[k] = 2 logdet[j], and both planes of the Green's-function
 * rows greens[p, j] go to row k of the same planes at greens_dst +
 * (greens - greens_src) (a buffer laid out like the one greens_src starts,
 * not overlapping it).  idx: afq_block_end_grouped's.
 */
int afq_reconfigure_gather_grouped(int count, const afq_system_step* sys, int total_walkers, const int* idx,
                                   const double* greens_src, double* greens_dst, void* stream);

/*
 * afq_energy_f32 (restricted walkers) for every system: local energies of the
 * Green's functions in `greens` into `energy` (meas/rhf.py:36-50).  Every
 * system of the table must have ceil(o / 16) == nt (1 or 2); max_n the largest n.
 */
int afq_energy_f32_grouped(int count, const afq_system_step* sys, int total_walkers, int max_n,
                           int nt, void* stream);

/*
 * afq_energy_bf16x3 (restricted walkers) for every system of a table whose
 * systems all have ceil(o / 16) == nt (1 or 2): wpb (1 or 2) walkers per
 * workgroup, tblock0 the prefix sums of ceil(nw / wpb), total_blocks their
 * sum; rot_chol the split planes of the half-rotated Cholesky (hi, then lo;
 * ldr = n rounded up to 32); max_n the largest n (<= 256) (meas/rhf.py:36-50).
 */
int afq_energy_bf16x3_grouped(int count, const afq_system_step* sys, int total_blocks, int max_n, int nt,
                              int wpb, void* stream);

/* 1 when afq_energy_bf16x3_grouped's LDS for (max_n, nt, wpb) fits in a CU,
 * else 0.  Host-only query, no device work; no counterpart in the reference. */
int afq_energy_bf16x3_grouped_fits(int max_n, int nt, int wpb);

/*
 * Local energies with the exchange as a quadratic form: replaces
 * energy_kernel_r / energy_kernel_u (meas/rhf.py:36-50, 75-105), the sum over g
 * of tr(f_g f_g) regrouped as vec(G_s)^T Y_s vec(G_s) with the fixed
 * Y_s[(a,q),(b,p)] = sum_g R_s[g,a,p] R_s[g,b,q].  greens: planes (2, nw,
 * nspin*on) (on = o*n); z: per sector s the planes (2, nw, on) of G_s Y_s at z
 * + s*2*nw*on (the caller's GEMM); c: planes (2, nw, naux), c_scale times the
 * force-bias contraction summed over sectors; rot_h1: (nspin, on).  energy:
 * complex128 (nw) = h0 + 2 e1 + 2 sum c^2 - exc (nspin 1) or
 * h0 + e1 + (sum c^2 - exc) / 2 (nspin 2), fp64 sums in a fixed order.
 */
int afq_energy_assemble(int nw, int on, int nspin, int naux, double h0, double c_scale, const double* greens,
                        const double* z, const double* c, const double* rot_h1, double* energy, void* stream);

/*
 * afq_energy_assemble for every system of a batch (restricted walkers): one
 * workgroup per walker (total_walkers = sum of nw); per system greens, exch_z,
 * c = fb / 2, rot_h1, h0 and energy from its descriptor.
 */
int afq_energy_assemble_grouped(int count, const afq_system_step* sys, int total_walkers, void* stream);

/*
 * afq_orthonormalize for every system: walkers_in -> walkers_out
 * (walkers.py:169-211).  max_n, max_o the largest n and o.
 */
int afq_orthonormalize_grouped(int count, const afq_system_step* sys, int total_walkers, int max_n,
                               int max_o, void* stream);

/*
 * ---------------------------------------------------------------------------
 * Hamiltonian ingestion: modified (pivoted, incomplete) Cholesky decomposition
 * of the two-electron integrals, replacing modified_cholesky
 * (integrals.py:3-46) as called by the front end (default.py:18-21).
 *   mat   : (size x size) float64, row pitch ldm: the 4-fold symmetric ERIs
 *           (pq|rs) in packed pair indices, size = n (n + 1) / 2, pair (m, p)
 *           with m >= p at m (m + 1) / 2 + p;
 *   vecs  : (max_vectors x size) float64, the Cholesky vectors v_0, v_1, ...;
 *   resid, partial : workspace of afq_mcd_workspace's sizes (doubles);
 *   state : float64[4] = {pivot value, pivot index, vectors so far, stopped}.
 * afq_mcd_start forms v_0 and the stop flag; afq_mcd_iterate enqueues
 * `iterations` iterations (each a no-op once stopped), so the caller can
 * enqueue a chunk, read state[3] and continue; the result is
 * v_0 .. v_{state[2]-1}.  Iteration: M += v_k^2, nu = argmax |diag - M|
 * (first index on ties), d = |diag - M|[nu],
 * v_{k+1} = (mat[nu] - sum_{i<=k} v_i[nu] v_i) / sqrt(d + 1e-10), while
 * |d| > max_error and fewer than max_vectors (<= size) vectors.
 * afq_unpack_pairs: chol[i][m][p] = vecs[i][pair(max(m,p), min(m,p))],
 * (nchol, n, n) float64 (integrals.py:39-45).
 * ---------------------------------------------------------------------------
 */
int afq_mcd_workspace(int size, long long* resid_doubles, long long* partial_doubles);
int afq_mcd_start(int size, const double* mat, long long ldm, double max_error, int max_vectors,
                  double* vecs, double* resid, double* partial, double* state, void* stream);
int afq_mcd_iterate(int size, const double* mat, long long ldm, double max_error, int max_vectors,
                    int iterations, double* vecs, double* resid, double* partial, double* state,
                    void* stream);
int afq_unpack_pairs(int nchol, int n, const double* vecs, long long ldv, double* chol, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* AFQMC_HIP_H */
