/*
 * trot_amd: MI355X (gfx950) kernels for the phaseless AFQMC walker step and
 * local-energy measurement of the `trot` package.
 *
 * C ABI.  Conventions:
 *   - all array arguments are device pointers; complex arrays are interleaved
 *     (re, im) doubles, i.e. the layout of complex128 / torch.complex128;
 *   - walkers are (W, N, n) complex, row-major (orbital row, occupied column);
 *   - trial orbitals are (N, n) real (the reference's RHF/UHF mo_coeff from
 *     a real SCF);
 *   - "green_t" is the transposed half-rotated Green's function
 *     GT = w (C^T w)^-1, (W, N, n) complex: G_half[p, j] = GT[j, p];
 *   - every function returns TROT_OK or a negative TROT_ERR_* code, never
 *     throws, and records a message retrievable with trot_last_error();
 *   - the caller owns all memory; scratch comes in as (ws, ws_bytes), sized
 *     by the matching *_workspace() query.  No entry point allocates;
 *   - `stream` is a hipStream_t (NULL: default stream).  Work is enqueued
 *     asynchronously on it.  Calls on one stream are ordered.
 *
 * Each entry point names the reference function (file:line, in the trot
 * repository) whose per-walker arithmetic it performs for all W walkers.
 */
#ifndef TROT_AMD_H
#define TROT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TROT_OK 0
#define TROT_ERR_INVALID (-1)     /* bad pointer or size */
#define TROT_ERR_UNSUPPORTED (-2) /* valid input outside what the kernels cover (e.g. nocc > 256) */
#define TROT_ERR_HIP (-3)         /* HIP runtime error (launch or configuration) */
#define TROT_ERR_WORKSPACE (-4)   /* workspace smaller than the *_workspace() query */

/* precision of the Cholesky-side MFMA GEMMs in the *_lp entry points */
#define TROT_PREC_F64 0    /* v_mfma_f64_16x16x4_f64 (the plain entry points) */
#define TROT_PREC_F32 1    /* v_mfma_f32_16x16x4_f32, FP32 accumulation */
#define TROT_PREC_BF16X3 2 /* 3-pass split BF16 (hi.hi + hi.lo + lo.hi) on v_mfma_f32_16x16x32_bf16 */

const char* trot_version(void);
const char* trot_last_error(void);

/* nocc: up to 256 per spin for the overlap / Green's function, force bias, local energy (FP64, FP32 and the
 * FP32-tile split-BF16 form), QR and the step; up to 64 for trot_local_energy_b3, trot_rot_chol_b3, the block
 * observables, the grouped inverse and the CISD entry points. */

/* Overlap matrix m = C^T w, det(m) and GT = w m^-1 for every walker.
 * Replaces: overlap_r / _det (trot/trial/rhf.py:35-47) for the determinant and
 * _half_green_from_overlap_matrix (trot/meas/rhf.py:21-25, called at
 * meas/rhf.py:121-122) for the Green's function.  green_t may be NULL. */
size_t trot_overlap_green_workspace(int W, int N, int n);
int trot_overlap_green(const double* trial, const double* walkers, int W, int N, int n, double* det,
                       double* green_t, void* ws, size_t ws_bytes, void* stream);

/* overlap[w] = det_a[w] * det_b[w]; det_b NULL gives det_a^2.
 * Replaces: overlap_r (trot/trial/rhf.py:45-47), overlap_u (rhf.py:50-54,
 * trot/trial/uhf.py:54-58). */
int trot_overlap_combine(const double* det_a, const double* det_b, double* overlap, int W, void* stream);

/* fb (+)= alpha * GT_flat . R'^T with R' (G, N*n) real, R'[g, j*n+p] = rot_chol[g, p, j].
 * Replaces: force_bias_kernel_rw_rh (trot/meas/rhf.py:119-125, alpha = 2),
 * force_bias_kernel_uw_rh (rhf.py:141-153; trot/meas/uhf.py:43-57: two calls, accumulate). */
int trot_force_bias(const double* green_t, const double* rot_chol_t, int W, int G, long Nn, double alpha,
                    int accumulate, double* fb, void* stream);

/* As trot_force_bias, with K = Nn split into slices when W x G alone does not fill the GPU (small walker sets,
 * fragments): FP64 partial products in ws (trot_force_bias_workspace(W, G, Nn) bytes, 0 when no split),
 * summed in slice order.  The slice count depends on (W, G, Nn) only.  Replaces: as trot_force_bias. */
size_t trot_force_bias_workspace(int W, int G, long Nn);
int trot_force_bias_splitk(const double* green_t, const double* rot_chol_t, int W, int G, long Nn, double alpha,
                           int accumulate, double* fb, void* ws, size_t ws_bytes, void* stream);

/* fs = -sqrt(dt)(i fb - mf); xbar = x - fs; shift_term = sum_g xbar mf; fb_term = sum_g (x fs - fs^2/2).
 * Replaces: trot/prop/afqmc.py:99-103.  fields (W, G) real; fb, xbar (W, G) complex; mf (G) complex. */
int trot_field_shift(const double* fields, const double* fb, const double* mf, double sqrt_dt, int W, int G,
                     double* xbar, double* shift_term, double* fb_term, void* stream);

/* V (W, N*N) complex = xbar (W, G) complex . chol_flat (G, N*N) real.
 * Replaces: _make_vhs_split_flat (trot/prop/chol_afqmc_ops.py:83-87). */
int trot_vhs_build(const double* xbar, const double* chol_flat, int W, int G, long N2, double* vhs, void* stream);

/* w_out[b] = expm (N, N real) . w_in[b]; not in place.
 * Replaces: _apply_one_body_half_array (trot/prop/chol_afqmc_ops.py:130-131). */
int trot_one_body_apply(const double* expm, const double* w_in, int W, int N, int n, double* w_out, void* stream);

/* w <- sum_{k=0..K} (i sqrt(dt) V)^k / k! w via term <- (i sqrt(dt)/k) V term; out += term.
 * Replaces: taylor_expm_action (trot/prop/utils.py:20-29) as used in
 * _apply_two_body_array (trot/prop/chol_afqmc_ops.py:152-162). */
size_t trot_taylor_workspace(int W, int N, int n);
int trot_taylor_apply(const double* vhs, double* w, int W, int N, int n, double sqrt_dt, int n_terms, void* ws,
                      size_t ws_bytes, void* stream);

/* Tile shape of the FP64 Taylor-term GEMMs of trot_taylor_apply (symblk 0) and trot_taylor_apply_symblk
 * (symblk 1) for one (W, N, n): 0 = 64x64 (default), 1 = 64x32, 2 = 32x64, 3 = 32x32 (orbitals x occupied).
 * Every shape gives bit-identical results; only the speed differs.  trot_autotune_taylor times one term in
 * each shape on scratch operands carved from ws (trot_autotune_taylor_workspace bytes), synchronizing the
 * stream (an error while the stream is being captured), and records the fastest in a process-wide table
 * guarded by a mutex; ms_out (4 floats) may be NULL.  No counterpart in the reference (XLA picks its own
 * tiles); used by the batched fragment form, whose walker sets have uneven small sizes. */
size_t trot_autotune_taylor_workspace(int W, int N, int n, int symblk);
int trot_autotune_taylor(int W, int N, int n, int symblk, void* ws, size_t ws_bytes, void* stream, int* shape_out,
                         float* ms_out);
int trot_set_taylor_tile(int W, int N, int n, int symblk, int shape);
int trot_get_taylor_tile(int W, int N, int n, int symblk);

/* In place: w_s <- E . taylor(i sqrt(dt), V(xbar), E . w_s, K) for one or two spin sets sharing V,
 * in walker chunks of `chunk` (<= 0: all).  walkers_b NULL for restricted walkers.
 * Replaces: _apply_trotter_r / _apply_trotter_u (trot/prop/chol_afqmc_ops.py:198-222). */
size_t trot_apply_trotter_workspace(int W, int N, int n_max, int chunk);
int trot_apply_trotter(double* walkers_a, int na, double* walkers_b, int nb, const double* xbar,
                       const double* chol_flat, const double* expm, int W, int N, int G, double sqrt_dt,
                       int n_terms, int chunk, void* ws, size_t ws_bytes, void* stream);

/* Phaseless importance, floor/cap masking, node count and population-control shift.
 * weights in/out (W real); scalars[0] = pop_control_ene_shift (in/out), scalars[1] = e_estimate;
 * nodes[0] += node encounters.  Replaces: trot/prop/afqmc.py:108-134. */
int trot_phaseless_update(double* weights, const double* ov_old, const double* ov_new, const double* shift_term,
                          const double* fb_term, double* scalars, long long* nodes, int W, double dt, double h0p_re,
                          double h0p_im, double w_floor, double w_cap, double damping, void* stream);

/* E_L per walker (complex) from cached GT.  rot_chol_s (G, n_s, N) real, rot_h1t_s (N, n_s) real.
 * Second spin NULL: RHF trial with restricted walkers (factor 2 for both spins).
 * Replaces: energy_kernel_rw_rh / _two_body_energy_restricted (trot/meas/rhf.py:41-60, 128-138),
 * energy_kernel_uw_rh (rhf.py:156-172, 63-98) and trot/meas/uhf.py:205-231. */
size_t trot_local_energy_workspace(int W, int G);
int trot_local_energy(const double* rot_chol_a, const double* rot_chol_b, const double* green_t_a,
                      const double* green_t_b, const double* rot_h1t_a, const double* rot_h1t_b, int W, int N,
                      int na, int nb, int G, double h0, double* energy, void* ws, size_t ws_bytes, void* stream);

/* Block energy tail: Re E_L clamped to e_estimate when non-finite or |E - e_est| > sqrt(2/dt); non-finite
 * walkers get weight 0 (weights in/out); out[0] = weighted mean (e_est if sum w = 0), out[1] = sum w;
 * scalars[1] (e_estimate) <- (1 - shift_ema) e_est + shift_ema out[0].
 * Replaces: trot/prop/blocks.py:236-257. */
int trot_block_energy(const double* energy, double* weights, double* scalars, double* out, int W, double dt,
                      double shift_ema, void* stream);

/* Block observables.  rdm1: out[w, spin] (out is (W, n_spin_out, N, N) complex) = GT[w] C^T, the mixed
 * one-body density matrix w (C^T w)^-1 C^T of a real trial C (N x n) and the walker's transposed half
 * Green's function GT (W, N, n) complex.
 * Replaces: trot/meas/rhf.py:191-216 (rdm1_kernel_rw / _uw), trot/meas/uhf.py:97-110 (rdm1_kernel_uw). */
int trot_rdm1_from_green(const double* green_t, const double* trial, int W, int N, int n, int n_spin_out, int spin,
                         double* out, void* stream);

/* Density correlations from dm (W, 2, N, N) complex: out (W, 3, N, N) = (uu, ud, dd) with
 * uu_ij = na_i na_j - ga_ij ga_ji + delta_ij na_i, ud_ij = na_i nb_j, n = diag(g).
 * Replaces: trot/meas/uhf.py:129-144 (_density_corr_from_greens), rhf.py:237-261. */
int trot_density_corr(const double* dm, int W, int N, double* out, void* stream);

/* Weighted block mean of per-walker samples x (W, L) doubles (complex samples as 2 doubles each):
 * out[l] = sum_w weights[w] x[w, l] / *wsum, or 0 when *wsum == 0 (wsum is a device scalar, e.g. out[1] of
 * trot_block_energy).  Fixed summation order: the same bits on every call.
 * Replaces: trot/prop/blocks.py:259-268. */
size_t trot_weighted_mean_workspace(int W, long L);
int trot_weighted_mean(const double* x, const double* weights, const double* wsum, int W, long L, double* out,
                       void* ws, size_t ws_bytes, void* stream);

/* In place Q factor of w with real positive diag(R) (two Cholesky-QR passes).  A walker whose Cholesky breaks
 * down (a pivot not finite or below 1e-8 of its Gram diagonal: ill-conditioned, kappa above ~1e4..1e8) is redone
 * from its original columns by Gram-Schmidt with reorthogonalisation.
 * Replaces: _qr / orthonormalize (trot/walkers.py:137-179). */
size_t trot_orthonormalize_workspace(int W, int N, int n);
int trot_orthonormalize(double* walkers, int W, int N, int n, void* ws, size_t ws_bytes, void* stream);

/* trot_orthonormalize that also keeps the norm: det_r[w] = prod diag(R) with diag(R) real positive, the
 * product of the two CholQR passes' Cholesky diagonals (workspace trot_orthonormalize_workspace).
 * Replaces: _qr / orthogonalize (trot/walkers.py:137-172), used by block_fp (trot/prop/blocks.py:451-458). */
int trot_orthogonalize(double* walkers, int W, int N, int n, double* det_r, void* ws, size_t ws_bytes, void* stream);

/* det_r[w] = prod_k L_kk for the Cholesky factor L of w^H w: |det R| of w = QR, which is the phase-fixed
 * diag(R) product of the reference (LAPACK's geqrf returns a real diagonal, so d^2/|d| = |d|).
 * Replaces: qr_norm (trot/walkers.py:182-205) per spin. */
size_t trot_qr_norm_workspace(int W, int N, int n);
int trot_qr_norm(const double* walkers, int W, int N, int n, double* det_r, void* ws, size_t ws_bytes, void* stream);

/* ---- free projection (trot/prop/afqmc_fp.py:18-72, trot/prop/blocks.py:415-490) ----
 * trot_fp_prepare: per walker z = -sqrt(dt) sum_g x_g mf_g + dt h0p, cst = exp(z) (optional),
 *   scale = cst^(1/n_total) (principal power), xbar = x + 0i (optional, (W, G) complex128).
 *   Replaces: shift_term / constants (afqmc_fp.py:95-96) and multiply_constants (trot/walkers.py:208-226).
 * trot_scale_rows: a[b, :] *= scale[b] for rows of row_complex complex numbers.
 * trot_fp_weights: norm = det_a^2 (det_b NULL, restricted walkers) or det_a det_b; weights_new = weights norm
 *   (afqmc_fp.py:107-108).
 * trot_fp_unnorm: weights[i] /= norms[idx[i]] after the comb (afqmc_fp.py:113-119).
 * trot_block_energy_fp: E_L -> ene0 where |E_L - ene0| > sqrt(2/dt); out[8] = complex pairs
 *   [sum w ov E / sum w ov, sum w ov, sum ov, sum |ov|] (blocks.py:466-488). */
int trot_fp_prepare(const double* fields, const double* mf, int W, int G, double sqrt_dt, double dt, double h0p_re,
                    double h0p_im, int n_total, double* xbar, double* cst, double* scale, void* stream);
int trot_scale_rows(double* a, const double* scale, int W, long row_complex, void* stream);
int trot_fp_weights(const double* weights, const double* det_a, const double* det_b, int W, double* norms,
                    double* weights_new, void* stream);
int trot_fp_unnorm(double* weights, const double* norms, const int* idx, int W, void* stream);
int trot_block_energy_fp(const double* energy, const double* weights, const double* overlaps, int W, double dt,
                         double ene0, double* out, void* stream);

/* Comb reconfiguration: idx[i] = searchsorted(cumsum|w|, tot (i + zeta)/W, 'left'), weights_new = tot/W.
 * Replaces: _sr_indices / stochastic_reconfiguration (trot/walkers.py:249-293). */
size_t trot_sr_workspace(int W);
int trot_sr_comb(const double* weights, double zeta, int W, int* idx, double* weights_new, void* ws,
                 size_t ws_bytes, void* stream);

/* trot_sr_comb with the comb offset zeta read from device memory (one double), so that the block
 * tail can run inside a captured HIP graph with zeta drawn on the device (trot_philox_uniform).
 * Replaces: _sr_indices with the uniform of trot/prop/blocks.py:270-271 (trot/walkers.py:249-254). */
int trot_sr_comb_dev(const double* weights, const double* zeta, int W, int* idx, double* weights_new, void* ws,
                     size_t ws_bytes, void* stream);

/* trot_sr_comb_dev without a workspace argument: the cumulative weights go to a static device array of
 * TROT_SR_NOALLOC_MAX_W doubles (2 MiB) that each GPU holds until the process ends.  Allocates nothing and
 * calls no allocating routine; W above the limit returns TROT_ERR_UNSUPPORTED.  Calls on one GPU share the
 * array, so calls that overlap on one GPU (two streams, two threads) race; calls ordered on one stream do not.
 * Results are bit-identical to trot_sr_comb_dev.
 * No counterpart in the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). */
#define TROT_SR_NOALLOC_MAX_W 262144
int trot_sr_comb_dev_noalloc(const double* weights, const double* zeta, int W, int* idx, double* weights_new,
                             void* stream);

/* Counter-based random numbers: Philox4x64-10 (Salmon et al., SC'11; numpy.random.Philox), block j of
 * draw `counter` for purpose p is Philox(counter words (j, counter, p, 0), key words (seed, 0)).
 * `state`, when non-null, is a device array {seed, counter} read in place of the by-value pair, so
 * a captured graph draws new numbers on every replay; trot_rng_advance adds `by` to state[1].
 * trot_philox_normal: n N(0,1) values, Box-Muller on the 53-bit uniforms of (x0, x1), (x2, x3) of block
 *   j -> values 4j..4j+3 (purpose 0).  Replaces: the threefry field draw of trot/prop/afqmc.py:91-93.
 * trot_philox_uniform: one uniform in [0, 1), (x0 >> 11) 2^-53 of block 0, purpose 1.
 *   Replaces: the comb offset draw of trot/prop/blocks.py:270-271.
 * trot_philox_raw: the raw 64-bit words of blocks 0..n_blocks-1 (4 per block), for checking. */
int trot_philox_raw(uint64_t seed, uint64_t counter, uint64_t purpose, long n_blocks, uint64_t* out, void* stream);
int trot_philox_normal(const uint64_t* state, uint64_t seed, uint64_t counter, long n, double* out, void* stream);
int trot_philox_uniform(const uint64_t* state, uint64_t seed, uint64_t counter, double* out, void* stream);
int trot_rng_advance(uint64_t* state, uint64_t by, void* stream);

/* dst[i, :] = src[idx[i], :] for rows of row_doubles doubles (walkers, Green's functions, overlaps).
 * Replaces: the gather w[idx] of trot/walkers.py:272-278. */
int trot_gather_rows(const double* src, double* dst, const int* idx, int W, long row_doubles, void* stream);

/* rot_chol[g] = C^T L_g (G, n, N) and rot_chol_t[g] = L_g^T C (G, N, n); either output may be NULL.
 * Replaces: build_meas_ctx (trot/meas/rhf.py:308-319; trot/meas/uhf.py:318-336). */
int trot_build_rot_chol(const double* chol, const double* trial, int G, int N, int n, double* rot_chol,
                        double* rot_chol_t, void* stream);

/* mf[g] = i tr(L_g dm) (complex), mfl = sum_g mf_g L_g (N x N complex), v0m = 1/2 sum_g L_g L_g (real).
 * Replaces: _mf_shifts and the contractions of _get_h1_eff
 * (trot/prop/chol_afqmc_ops.py:74-76, 90-100). */
size_t trot_build_prop_workspace(int G, int N);
int trot_build_prop_tensors(const double* chol, const double* dm, int G, int N, double* mf, double* mfl,
                            double* v0m, void* ws, size_t ws_bytes, void* stream);

/* ---- symmetric Cholesky vectors (every L_g == L_g^T, as for real-orbital ERIs) ----
 * V = sum_g xbar_g L_g is then symmetric.  "sym-block" storage keeps only the 8 x 8 blocks (bi <= bj) of a
 * symmetric N x N matrix, nb = ceil(N/8) blocks per side, block-row major, each block row-major:
 * trot_symblk_cols(N) = 64 nb (nb + 1) / 2 numbers per matrix.  trot_vhs_build(_lp) with chol_flat in this
 * storage (N2 = trot_symblk_cols(N)) produces V in this storage at ~half the flops and bytes, and the
 * *_symblk Taylor entry points read it.  Same result as the dense path up to summation order.
 * Replaces: _make_vhs_split_flat + taylor_expm_action (trot/prop/chol_afqmc_ops.py:83-87, 152-162;
 * trot/prop/utils.py:20-29) for symmetric chol. */
long trot_symblk_cols(int N);
int trot_pack_symblk(const double* chol, int G, int N, void* out, int out_f32, void* stream);
int trot_taylor_apply_symblk(const double* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt, int n_terms,
                             void* ws, size_t ws_bytes, void* stream);
int trot_taylor_apply_lp_symblk(const float* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt,
                                int n_terms, int prec, void* ws, size_t ws_bytes, void* stream);

/* ---- 3-pass split-BF16 NT GEMM on pre-split operands ----
 * An operand matrix X (rows x K) is stored as bf16 planes x = hi + lo: [re_hi, re_lo, im_hi, im_lo] for
 * complex, [hi, lo] for real, each plane (trot_b3_rows_pad(rows), trot_b3_k_pad(K)), K contiguous, zero
 * padded.  trot_gemm_b3 computes C = alpha A B^T (+ beta C) with A (M x K, real or complex) and B
 * (N x K, real) in that storage on v_mfma_f32_16x16x32_bf16 (lo.hi + hi.lo + hi.hi, FP32 accumulation).
 * Used for the VHS (A = xbar, B = chol in sym-block or dense layout, transposed) and the force bias
 * (A = GT, B = R'): the Cholesky GEMMs of trot/prop/chol_afqmc_ops.py:83-87 and trot/meas/rhf.py:125. */
long trot_b3_rows_pad(int rows);
long trot_b3_k_pad(int K);
size_t trot_b3_planes_bytes(int rows, int K, int is_complex);
int trot_split_bf16(const void* src, int src_f32, int is_complex, int transposed, long ld, int rows, int K,
                    void* dst, void* stream);
size_t trot_gemm_b3_workspace(int M, int N, int a_complex, int ksplit);
/* c_f64: 0 FP32 (complex64) output, 1 FP64 (complex128), 2 split complex64 for trot_trotter_b3v (complex A,
 * beta 0, no split K): two planes of M x ldc words bf16(re) | bf16(im) << 16, hi then lo. */
int trot_gemm_b3(const void* a_planes, int a_complex, const void* b_planes, int M, int N, int K, void* C, int c_f64,
                 long ldc, double alpha_re, double alpha_im, double beta, int ksplit, void* ws, size_t ws_bytes,
                 void* stream);

/* Taylor series with every product on split-BF16 MFMA: V (complex64, sym-block storage,
 * (W, trot_symblk_cols(N))) is read as stored and split into bf16 hi/lo planes in LDS; the terms travel
 * as split planes of T^T, are summed in FP32 and added to w (complex128, in place) in FP64.
 * Workspace trot_taylor_b3_workspace(W, N, n).
 * Replaces: taylor_expm_action (trot/prop/utils.py:20-29) in _apply_two_body_array
 * (trot/prop/chol_afqmc_ops.py:152-162). */
size_t trot_taylor_b3_workspace(int W, int N, int n);
int trot_taylor_apply_b3_symblk(const float* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt, int n_terms,
                                void* ws, size_t ws_bytes, void* stream);

/* (E - I) for E = exp(-dt/2 h1_eff) (N x N, FP64) as split bf16 planes, trot_onebody_b3_planes_bytes(N) bytes;
 * built once per propagation context (exp_h1_half, trot/prop/chol_afqmc_ops.py:116). */
size_t trot_onebody_b3_planes_bytes(int N);
int trot_onebody_b3_planes(const double* expm, int N, void* planes, void* stream);

/* One spin set: w_out <- E . taylor(i sqrt(dt), V, E . w_in, K) with every product on split-BF16 MFMA; each
 * one-body factor is applied as w + (E - I) w with the identity term exact in FP64.  w_out may equal w_in.
 * Workspace trot_taylor_b3_workspace(W, N, n).
 * Replaces: the per-spin body of _apply_trotter_r / _apply_trotter_u (trot/prop/chol_afqmc_ops.py:198-222):
 * _apply_one_body_half_array (:130-131), _apply_two_body_array (:152-162), _apply_one_body_half_array. */
int trot_trotter_b3_symblk(const float* vhs_symblk, const void* onebody_planes, const double* w_in, double* w_out,
                           int W, int N, int n, double sqrt_dt, int n_terms, void* ws, size_t ws_bytes, void* stream);

/* trot_trotter_b3_symblk with V as split planes: v_planes = [hi | lo], each (W, trot_symblk_cols(N)) 32-bit
 * words bf16(re) | bf16(im) << 16 with lo = bf16(x - hi), as trot_gemm_b3 writes with c_f64 = 2 (the
 * same bf16 operands the complex64 form makes in LDS, so hi_from > n_terms gives identical results).
 * Taylor terms k >= hi_from are one bf16 pass on the hi planes of V and of the term (half the V and term
 * bytes, a quarter of the products): term k is smaller than the first by about (sqrt(dt) |V|)^(k-1) / k!,
 * so from k = 3 on a 2^-8 relative product error stays at the size of the split-BF16 error of term 1.
 * Replaces: as trot_trotter_b3_symblk (trot/prop/chol_afqmc_ops.py:198-222, trot/prop/utils.py:20-29). */
int trot_trotter_b3v(const void* v_planes, int hi_from, const void* onebody_planes, const double* w_in,
                     double* w_out, int W, int N, int n, double sqrt_dt, int n_terms, void* ws, size_t ws_bytes,
                     void* stream);

/* The launches of trot_trotter_b3v as descriptors, for grouping them over many walker sets (the batching
 * form's fragments in mixed precision).  trot_trotter_b3v_stages fills, in launch order, the kind
 * (TROT_B3_STAGE_*), the workgroup count and the descriptor (trot_prop_b3_stage_bytes() bytes) of each launch
 * the same call of trot_trotter_b3v would make, and returns their number (nothing is launched; ws stays in
 * use).  trot_prop_b3_grouped runs one kind for nprob sets from a device copy of their descriptors and of the
 * nprob + 1 prefix sums of their workgroup counts; each workgroup runs exactly the body it runs in its own
 * set's launch, so a grouped sequence is bit-identical to trot_trotter_b3v per set.  TROT_B3_STAGE_COPY has
 * no grouped form (TROT_ERR_UNSUPPORTED).
 * Replaces: _apply_trotter_u per fragment (trot/prop/chol_afqmc_ops.py:211-222), the reference stepping
 * fragments one at a time (trot/lno.py:19-141). */
/* Grouped forms of the split-BF16 operand split (complex FP64, row-major), the wide split-BF16 GEMM of
 * trot_gemm_b3 with complex A (its K slices included) and its slice reduction, for the batching form's force
 * bias and VHS in mixed precision: *_problem fills one descriptor (*_problem_bytes() bytes) with the arguments
 * of one call and returns its workgroup count (trot_b3_reduce_problem: 0 for a product without split K);
 * *_grouped runs nprob problems from device copies of the descriptors and of the nprob + 1 prefix sums of
 * their workgroup counts, each workgroup exactly as in its own call's launch (bit-identical).
 * trot_gemm_b3_grouped runs problems of one output kind (c_kind as trot_gemm_b3's c_f64).
 * Replaces: the per-fragment force bias and VHS (trot/meas/uhf.py:43-57, trot/prop/chol_afqmc_ops.py:83-87,
 * 247-257) of fragments stepped one at a time (trot/lno.py:19-141). */
size_t trot_split_bf16_problem_bytes(void);
long trot_split_bf16_problem(const void* src, long ld, int rows, int K, void* dst, void* desc);
int trot_split_bf16_grouped(const void* probs, const int* start, int nprob, int total_wg, void* stream);
size_t trot_gemm_b3_problem_bytes(void);
long trot_gemm_b3_problem(const void* a_planes, const void* b_planes, int M, int N, int K, void* C, int c_kind,
                          long ldc, double alpha_re, double alpha_im, double beta, int ksplit, void* ws,
                          size_t ws_bytes, void* desc);
int trot_gemm_b3_grouped(int c_kind, const void* probs, const int* start, int nprob, int total_wg, void* stream);
size_t trot_b3_reduce_problem_bytes(void);
long trot_b3_reduce_problem(const void* gemm_desc, void* desc);
int trot_b3_reduce_grouped(const void* probs, const int* start, int nprob, int total_wg, void* stream);

#define TROT_B3_STAGE_SPLIT 0
#define TROT_B3_STAGE_ONEBODY 1
#define TROT_B3_STAGE_TERM_C64 2
#define TROT_B3_STAGE_TERM3 3
#define TROT_B3_STAGE_TERM1 4
#define TROT_B3_STAGE_COPY 5
size_t trot_prop_b3_stage_bytes(void);
long trot_trotter_b3v_stages(const void* v_planes, int hi_from, const void* onebody_planes, const double* w_in,
                             double* w_out, int W, int N, int n, double sqrt_dt, int n_terms, void* ws,
                             size_t ws_bytes, int max_stages, int* kinds, long* workgroups, void* descs);
int trot_prop_b3_grouped(int kind, const void* probs, const int* start, int nprob, int total_wg, void* stream);

/* Split-BF16 planes of rot_chol (G, n, N) FP64 for trot_local_energy_b3, trot_rot_chol_b3_bytes(G, n, N) bytes;
 * made once per measurement context (trot/meas/rhf.py:308-319 builds rot_chol). */
size_t trot_rot_chol_b3_bytes(int G, int n, int N);
int trot_rot_chol_b3(const double* rot_chol, int G, int n, int N, void* planes, void* stream);

/* E_L per walker as trot_local_energy, with the exchange contractions f_g = rot_chol[g] . GT on split-BF16
 * MFMA (FP32 accumulation) for all walkers and Cholesky vectors in one batched GEMM; traces, exchange sums and
 * the sums over g in FP64.  rc_planes_s from trot_rot_chol_b3.  Second spin NULL: restricted walkers with an
 * RHF trial.  Workspace trot_local_energy_b3_workspace(W, N, G, second spin present).
 * Replaces: energy_kernel_rw_rh / _two_body_energy_restricted (trot/meas/rhf.py:41-60, 128-138),
 * energy_kernel_uw_rh (rhf.py:156-172, 63-98) and trot/meas/uhf.py:205-231. */
size_t trot_local_energy_b3_workspace(int W, int N, int G, int two_spins);
int trot_local_energy_b3(const void* rc_planes_a, const void* rc_planes_b, const double* green_t_a,
                         const double* green_t_b, const double* rot_h1t_a, const double* rot_h1t_b, int W, int N,
                         int na, int nb, int G, double h0, double* energy, void* ws, size_t ws_bytes, void* stream);

/* ---- reduced-precision forms (prec = TROT_PREC_F32 or TROT_PREC_BF16X3) ----
 * Cholesky-side operands are FP32 copies made at context build; FP64 walker-side operands are rounded
 * (and split) as they are staged into LDS; accumulation is FP32 on MFMA.  The reference's own mixed mode
 * (trot/prop/chol_afqmc_ops.py:247-257; runtime_layout.py:167) casts chol_flat to float32 and xbar to
 * complex64 for the VHS GEMM. */

/* As trot_vhs_build with chol_flat FP32; vhs is complex64 (vhs_f64 = 0) or complex128 (vhs_f64 = 1).
 * Replaces: _make_vhs_split_flat in mixed precision (trot/prop/chol_afqmc_ops.py:83-87, 247-257). */
int trot_vhs_build_lp(const double* xbar, const float* chol_flat, int W, int G, long N2, void* vhs, int vhs_f64,
                      int prec, void* stream);

/* As trot_force_bias with R' FP32; fb stays complex128.
 * Replaces: force_bias_kernel_rw_rh / _uw_rh (trot/meas/rhf.py:119-153; trot/meas/uhf.py:43-57). */
int trot_force_bias_lp(const double* green_t, const float* rot_chol_t, int W, int G, long Nn, double alpha,
                       int accumulate, double* fb, int prec, void* stream);

/* As trot_one_body_apply with the propagator in FP32 (FP32 or split-BF16 MFMA); walkers stay complex128.
 * Replaces: _apply_one_body_half_array (trot/prop/chol_afqmc_ops.py:130-131). */
int trot_one_body_apply_lp(const float* expm32, const double* w_in, int W, int N, int n, double* w_out, int prec,
                           void* stream);

/* As trot_taylor_apply with V complex64 and complex64 terms; the running sum w stays complex128.
 * Replaces: taylor_expm_action (trot/prop/utils.py:20-29). */
size_t trot_taylor_lp_workspace(int W, int N, int n);
int trot_taylor_apply_lp(const float* vhs, double* w, int W, int N, int n, double sqrt_dt, int n_terms, int prec,
                         void* ws, size_t ws_bytes, void* stream);

/* As trot_local_energy with rot_chol FP32 (exchange GEMM on MFMA FP32 / split BF16; sums in FP64).
 * Replaces: trot/meas/rhf.py:41-172; trot/meas/uhf.py:205-231. */
int trot_local_energy_lp(const float* rot_chol_a, const float* rot_chol_b, const double* green_t_a,
                         const double* green_t_b, const double* rot_h1t_a, const double* rot_h1t_b, int W, int N,
                         int na, int nb, int G, double h0, double* energy, int prec, void* ws, size_t ws_bytes,
                         void* stream);

/* ---- grouped launches: one launch per operation over many walker sets (the batching form) ----
 * Each operation of the FP64 step over many independent walker sets (LNO-style fragments) runs as
 * one launch: trot_gemm_problem describes one walker set's problem (the same operands, strides and K
 * order as the operation's own entry point: trot_one_body_apply, one term of trot_taylor_apply_symblk,
 * the two GEMMs of trot_overlap_green, trot_vhs_build) and returns its workgroup count; the caller
 * gathers the descriptors of all sets into device memory with prefix sums of the counts and launches
 * trot_gemm_f64_grouped.  Every workgroup computes exactly what it would in a launch of its own set, so
 * results are bit-identical to launching the sets one at a time.  trot_inv_problem / trot_inv_det_grouped
 * do the same for the inverse and determinant of trot_overlap_green, trot_copy_problem /
 * trot_copy_grouped for device-to-device copies (sizes multiples of 16 bytes).
 * Replaces: running trot/prop/afqmc.py:82-144 once per fragment (the reference batches nothing across
 * fragments; its LNO driver loops over them). */
#define TROT_GOP_ONE_BODY 0
#define TROT_GOP_TAYLOR 1
#define TROT_GOP_OVERLAP 2
#define TROT_GOP_GREEN 3
#define TROT_GOP_VHS 4
#define TROT_GOP_TAYLOR_NARROW 5
#define TROT_GOP_FORCE_BIAS 6 /* problems from trot_force_bias_problem, not trot_gemm_problem */
#define TROT_GOP_TAYLOR_64x64 7    /* TROT_GOP_TAYLOR on 64 x 64 tiles (same bits; the tile shapes of */
#define TROT_GOP_TAYLOR_64x32 8    /* trot_autotune_taylor: 64x64, 64x32, TAYLOR = 32x64, NARROW = 32x32) */
#define TROT_GOP_ONE_BODY_NARROW 9 /* TROT_GOP_ONE_BODY on 64 x 32 tiles (n <= 32; same bits) */
size_t trot_gemm_problem_bytes(void);
long trot_gemm_problem(int op, const double* a, const double* b, double* c, double* d, int W, int N, int n,
                       double coef, void* desc);
int trot_gemm_f64_grouped(int op, const void* probs, const int* start, int nprob, int total_wg, void* stream);
size_t trot_inv_problem_bytes(void);
long trot_inv_problem(const double* m, double* minv, double* det, int W, int n, void* desc);
int trot_inv_det_grouped(const void* probs, const int* start, int nprob, int total_w, void* stream);
size_t trot_copy_problem_bytes(void);
int trot_copy_problem(const void* src, void* dst, long bytes, void* desc);
int trot_copy_grouped(const void* probs, int nprob, long max_bytes, void* stream);
/* The rest of the step in the same form.  trot_force_bias_problem: trot_force_bias_splitk's GEMM as one
 * grouped problem (op TROT_GOP_FORCE_BIAS); with trot_force_bias_slices(W, G, Nn) > 1 slices `out` takes
 * the slices' partial products and a trot_sum_problem adds them in slice order into fb.
 * trot_shift_problem / trot_shift_grouped: trot_field_shift per set (trot/prop/afqmc.py:99-103).
 * trot_update_problem / trot_update_grouped: trot_overlap_combine, trot_phaseless_update and the copy of
 * the new overlaps over the old ones, one workgroup per set (afqmc.py:108-134).  All bit-identical to the
 * per-set entry points. */
int trot_force_bias_slices(int W, int G, long Nn);
long trot_force_bias_problem(const double* green_t, const double* rot_chol_t, int W, int G, long Nn, double alpha,
                             int accumulate, double* out, void* desc);
size_t trot_sum_problem_bytes(void);
long trot_sum_problem(const double* src, double* dst, int nb, long len, double scale, int accumulate, void* desc);
int trot_sum_grouped(const void* probs, const int* start, int nprob, int total_blocks, void* stream);
size_t trot_shift_problem_bytes(void);
long trot_shift_problem(const double* fields, const double* fb, const double* mf, double sqrt_dt, int W, int G,
                        double* xbar, double* shift_term, double* fb_term, void* desc);
int trot_shift_grouped(const void* probs, const int* start, int nprob, int total_w, void* stream);
size_t trot_update_problem_bytes(void);
int trot_update_problem(const double* det_a, const double* det_b, double* overlaps, const double* shift_term,
                        const double* fb_term, double* weights, double* scalars, long long* nodes, int W, double dt,
                        double h0p_re, double h0p_im, double w_floor, double w_cap, double damping, void* desc);
int trot_update_grouped(const void* probs, int nprob, void* stream);

/* ---- building blocks of the hand-written adjoint (reverse mode) of the step and the block energy ----
 * trot_gemm_batched: C[b] = alpha op(A[b]) op(B[b]) + beta C[b] for b < batch on FP64 MFMA, op TROT_OP_N (as
 *   stored), TROT_OP_T (transpose), TROT_OP_C (conjugate transpose), TROT_OP_R (conjugate); A is M x K and B
 *   K x N after op; real or interleaved-complex operands (conjugation of a real operand is an error), complex C
 *   whenever either operand is complex; leading dimensions and batch strides in elements (a stride of 0
 *   shares the operand), alpha complex, beta real.
 * trot_inverse_det: minv[b] = m[b]^-1 and det[b] = det m[b] (complex n x n, n <= 256, m != minv), the
 *   kernels of trot_overlap_green (register-resident Gauss-Jordan up to 64, LDS-resident up to 96, in place
 *   in minv above).
 * Replaces: the matrix products and the inverse that jax.grad takes through afqmc_step and the local energy
 * (trot/prop/afqmc.py:79-144, trot/prop/utils.py:20-29, trot/meas/rhf.py:21-25, 128-138; the reference is
 * differentiable end to end, README.md:11). */
#define TROT_OP_N 0
#define TROT_OP_T 1
#define TROT_OP_C 2
#define TROT_OP_R 3
int trot_gemm_batched(int opa, int opb, int a_complex, int b_complex, int M, int N, int K, double alpha_re,
                      double alpha_im, const double* A, long lda, long sA, const double* B, long ldb, long sB,
                      double beta, double* C, long ldc, long sC, int batch, void* stream);
int trot_inverse_det(const double* m, int W, int n, double* minv, double* det, void* stream);

/* The whole FP64 propagation of one small walker set as one grouped problem: w <- E . T_K(i sqrt(dt) V) . E w with
 * V in sym-block storage (W, trot_symblk_cols(N)) complex128 and E (N, N) real, w (W, N, n) in place, sum a
 * (W, N, n) scratch (the Taylor running sum).  trot_fused_prop_grouped runs one workgroup per walker of every
 * problem: the one-body product, the n_terms Taylor terms and the last one-body product with the current operand
 * resident in LDS, each output element accumulated by the same MFMA sequence as trot_one_body_apply and
 * trot_taylor_apply_symblk (bit-identical results).  trot_fused_prop_problem returns W, or TROT_ERR_UNSUPPORTED
 * when pad16(N) pad16(n) 16 bytes exceed 160 KiB, N > 256 or n > 48; max_lds is the largest of those byte counts
 * over the problems of a launch.
 * Replaces: _apply_trotter_r / _apply_trotter_u per walker set (trot/prop/chol_afqmc_ops.py:198-222) for the
 * batching form's fragments (the reference steps fragments one at a time). */
size_t trot_fused_prop_problem_bytes(void);
long trot_fused_prop_problem(const double* vsym, const double* expm, double* w, double* sum, int W, int N, int n,
                             double sqrt_dt, int n_terms, void* desc);
int trot_fused_prop_grouped(const void* probs, const int* start, int nprob, int total_wg, long max_lds, void* stream);

/* ---- CISD trial, restricted walkers (FP64) ----
 * The reference determinant occupies the first nocc orbitals; the active block is occupied orbitals
 * core .. nocc-1 (na = nocc - core) and virtuals nocc .. nocc+nvir-1; ci1 (na, nvir) and ci2 (na, nvir,
 * na, nvir) are real.  GT = w w[:nocc]^-1 and det = det(w[:nocc]) come from trot_overlap_green with the
 * trial I[:, :nocc].  Cholesky vectors and h1 must be symmetric (real orbitals).
 *
 * trot_cisd_doubles_matrix: D[(p,t),(q,u)] = 2 ci2[p,t,q,u] - ci2[p,u,q,t], (na nvir)^2 doubles: both
 *   doubles contractions of trot/trial/cisd.py:99-145 and trot/meas/cisd.py:179-198 in one matrix.
 * trot_pack_sym: upper triangles of `count` symmetric N x N matrices, row-major, N (N+1) / 2 each
 *   (the K index of the CISD force bias).
 * trot_cisd_overlap: Omega = 1 + 2 <ci1, x> + x^T D x, overlap = det^2 Omega (ovl may be NULL), and
 *   Bs = (ci1 + x D) / Omega (W, na, nvir).  Replaces trot/trial/cisd.py:148-166 (overlap_r).
 * trot_cisd_ps: Ps = greenp Bs^T (W, N, na), greenp of trot/meas/cisd.py:24-44.
 * trot_cisd_gamma: gamma = G0 - Ps G_act (W, N, N): the mixed 1-RDM of trot/meas/cisd.py:772-820
 *   (rdm1_kernel_rw, one spin); packed != 0 writes gamma[i,j] + gamma[j,i] (i < j) and gamma[i,i] in the
 *   trot_pack_sym order instead, so that the force bias of trot/meas/cisd.py:548-564
 *   (force_bias_kernel_rw_rh_high) is fb = 2 gamma_packed L_packed^T (trot_force_bias_splitk with
 *   Nn = N (N+1) / 2).
 * trot_cisd_energy: local energies (W complex) of trot/meas/cisd.py:608-769 (energy_kernel_rw_rh);
 *   lx = the G Cholesky vectors followed by h1, (G+1, N, N); walkers in chunks of wc, the workspace
 *   from trot_cisd_energy_workspace(wc, ..., b3).  d_planes NULL: the doubles term x_g^T D x_g in FP64;
 *   else split-BF16 planes of D^T (trot_split_bf16, transposed, nvir na rows and columns) and that
 *   term on split-BF16 MFMA (the reference computes it in float32 unless testing=True, :745-768). */
int trot_cisd_doubles_matrix(const double* ci2, int na, int nvir, double* D, void* stream);
int trot_pack_sym(const double* m, int count, int N, double* out, void* stream);
size_t trot_cisd_overlap_workspace(int W, int na, int nvir);
int trot_cisd_overlap(const double* green_t, const double* det, const double* ci1, const double* D, int W, int N,
                      int nocc, int core, int na, int nvir, double* omega, double* ovl, double* bs, void* ws,
                      size_t ws_bytes, void* stream);
int trot_cisd_ps(const double* green_t, const double* bs, int W, int N, int nocc, int core, int na, int nvir,
                 double* ps, void* stream);
int trot_cisd_gamma(const double* green_t, const double* ps, int W, int N, int nocc, int core, int na, int packed,
                    double* out, void* stream);
size_t trot_cisd_energy_workspace(int wc, int N, int nocc, int G, int na, int nvir, int b3);
int trot_cisd_energy(const double* green_t, const double* bs, const double* ps, const double* omega,
                     const double* lx, const double* D, const void* d_planes, int W, int N, int nocc, int G, int core,
                     int na, int nvir, double h0, int wc, double* e, void* ws, size_t ws_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TROT_AMD_H */
