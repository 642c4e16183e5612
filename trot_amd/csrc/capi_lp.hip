// C-ABI entry points of the reduced-precision forms (FP32 / split-BF16 MFMA).
//
// The Cholesky-side operands (chol_flat, rot_chol_t, rot_chol) are FP32
// copies made once at context build.  Walker-side operands stay FP64 in HBM
// and are rounded (and split, for BF16X3) on their way into LDS.  Outputs that
// feed FP64 arithmetic (force bias, energy, walkers) are FP64; the VHS and
// the Taylor terms may be FP32 (complex64).
#include "capi_util.hpp"
#include "common.hpp"
#include "gemm_lp.hpp"
#include "kernels.hpp"

static int mode_of(int prec) {
  if (prec == TROT_PREC_F32) return MODE_F32;
  if (prec == TROT_PREC_BF16X3) return MODE_BF16X3;
  return 0;
}

#define DISPATCH_MODE(prec, CALL)                                                      \
  do {                                                                                 \
    if (mode_of(prec) == MODE_F32) {                                                   \
      constexpr int MODE = MODE_F32;                                                   \
      CALL;                                                                            \
    } else if (mode_of(prec) == MODE_BF16X3) {                                         \
      constexpr int MODE = MODE_BF16X3;                                                \
      CALL;                                                                            \
    } else {                                                                           \
      return fail(TROT_ERR_INVALID, "unknown precision (TROT_PREC_F32 or _BF16X3)");   \
    }                                                                                  \
  } while (0)

extern "C" {

// V (W, N^2) complex (complex64, or complex128 if vhs_f64) = xbar (W, G) complex128 . chol_flat (G, N^2) f32
int trot_vhs_build_lp(const double* xbar, const float* chol_flat, int W, int G, long N2, void* vhs, int vhs_f64,
                      int prec, void* stream) {
  if (!xbar || !chol_flat || !vhs) return fail(TROT_ERR_INVALID, "vhs_build_lp: null pointer");
  if (W < 1 || G < 1 || N2 < 1 || N2 > 0x7fffffffL) return fail(TROT_ERR_INVALID, "vhs_build_lp: bad sizes");
  hipStream_t s = (hipStream_t)stream;
  if (vhs_f64) {
    GemmLPP<double, float, double> p{};
    p.A = xbar; p.lda = G; p.B = chol_flat; p.ldb = N2; p.C = (double*)vhs; p.ldc = N2;
    p.M = W; p.N = (int)N2; p.K = G; p.ar = 1.f;
    DISPATCH_MODE(prec, HIPCHK((launch_gemm_lp<64, 64, MODE, true, false, double, float, double, true, true, false, 0>(
                                p, 1, s))));
  } else {
    GemmLPP<double, float, float> p{};
    p.A = xbar; p.lda = G; p.B = chol_flat; p.ldb = N2; p.C = (float*)vhs; p.ldc = N2;
    p.M = W; p.N = (int)N2; p.K = G; p.ar = 1.f;
    DISPATCH_MODE(prec, HIPCHK((launch_gemm_lp<64, 64, MODE, true, false, double, float, float, true, true, false, 0>(
                                p, 1, s))));
  }
  return TROT_OK;
}

// fb (W, G) complex128 (+)= alpha GT_flat (W, Nn) complex128 . R'^T, R' (G, Nn) f32
int trot_force_bias_lp(const double* green_t, const float* rot_chol_t, int W, int G, long Nn, double alpha,
                       int accumulate, double* fb, int prec, void* stream) {
  if (!green_t || !rot_chol_t || !fb) return fail(TROT_ERR_INVALID, "force_bias_lp: null pointer");
  if (W < 1 || G < 1 || Nn < 1 || Nn > 0x7fffffffL) return fail(TROT_ERR_INVALID, "force_bias_lp: bad sizes");
  GemmLPP<double, float, double> p{};
  p.A = green_t; p.lda = Nn; p.B = rot_chol_t; p.ldb = Nn; p.C = fb; p.ldc = G;
  p.M = W; p.N = G; p.K = (int)Nn; p.ar = (float)alpha; p.beta = accumulate ? 1.0 : 0.0;
  DISPATCH_MODE(prec, HIPCHK((launch_gemm_lp<64, 64, MODE, true, false, double, float, double, true, false, false, 0>(
                              p, 1, (hipStream_t)stream))));
  return TROT_OK;
}

// w_out[b] (complex128) = expm32 (N x N, f32) . w_in[b] (complex128), FP32 / split-BF16 MFMA
int trot_one_body_apply_lp(const float* expm32, const double* w_in, int W, int N, int n, double* w_out, int prec,
                           void* stream) {
  if (!expm32 || !w_in || !w_out) return fail(TROT_ERR_INVALID, "one_body_apply_lp: null pointer");
  if (w_in == w_out) return fail(TROT_ERR_INVALID, "one_body_apply_lp: in-place not supported");
  if (W < 1 || N < 1 || n < 1) return fail(TROT_ERR_INVALID, "one_body_apply_lp: bad sizes");
  GemmLPP<float, double, double> p{};
  p.A = expm32; p.lda = N; p.sA = 0;
  p.B = w_in; p.ldb = n; p.sB = (long)N * n;
  p.C = w_out; p.ldc = n; p.sC = (long)N * n;
  p.M = N; p.N = n; p.K = N; p.ar = 1.f;
  DISPATCH_MODE(prec, HIPCHK((launch_gemm_lp<64, 64, MODE, false, true, float, double, double, true, true, false, 0>(
                              p, W, (hipStream_t)stream))));
  return TROT_OK;
}

size_t trot_taylor_lp_workspace(int W, int N, int n) { return 2 * align256((size_t)W * N * n * 8); }

// w (complex128, in/out) <- sum_k (i sqrt(dt) V)^k / k! w with V complex64 and the terms in complex64;
// the running sum stays complex128.
int trot_taylor_apply_lp(const float* vhs, double* w, int W, int N, int n, double sqrt_dt, int n_terms, int prec,
                         void* ws, size_t ws_bytes, void* stream) {
  if (!vhs || !w) return fail(TROT_ERR_INVALID, "taylor_apply_lp: null pointer");
  if (W < 1 || N < 1 || n < 1 || n_terms < 0) return fail(TROT_ERR_INVALID, "taylor_apply_lp: bad sizes");
  Carve c{(char*)ws, ws_bytes, 0};
  float* t0 = c.take<float>((size_t)W * N * n * 2);
  float* t1 = c.take<float>((size_t)W * N * n * 2);
  if (!t0 || !t1) return fail(TROT_ERR_WORKSPACE, "taylor_apply_lp: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n_terms == 0) return TROT_OK;
  // term 0 = w rounded to complex64 (the GEMM would round it on load anyway); w itself only accumulates
  HIPCHK(launch_c128_to_c64(w, t0, (long)W * N * n, s));
  float* prev = t0;
  float* next = t1;
  for (int k = 1; k <= n_terms; ++k) {
    GemmLPP<float, float, float> p{};
    p.A = vhs; p.lda = N; p.sA = (long)N * N;
    p.B = prev; p.ldb = n; p.sB = (long)N * n;
    p.C = next; p.ldc = n; p.sC = (long)N * n;
    p.C2 = w; p.ldc2 = n; p.sC2 = (long)N * n;
    p.M = N; p.N = n; p.K = N; p.ar = 0.f; p.ai = (float)(sqrt_dt / k);
    DISPATCH_MODE(prec, HIPCHK((launch_gemm_lp<64, 64, MODE, true, true, float, float, float, true, true, false, 1>(
                                p, W, s))));
    float* tmp = prev;
    prev = next;
    next = tmp;
  }
  return TROT_OK;
}

// As trot_taylor_apply_lp with V (complex64) in sym-block storage.
int trot_taylor_apply_lp_symblk(const float* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt, int n_terms,
                                int prec, void* ws, size_t ws_bytes, void* stream) {
  if (!vhs_symblk || !w) return fail(TROT_ERR_INVALID, "taylor_apply_lp_symblk: null pointer");
  if (W < 1 || N < 1 || n < 1 || n_terms < 0) return fail(TROT_ERR_INVALID, "taylor_apply_lp_symblk: bad sizes");
  Carve c{(char*)ws, ws_bytes, 0};
  float* t0 = c.take<float>((size_t)W * N * n * 2);
  float* t1 = c.take<float>((size_t)W * N * n * 2);
  if (!t0 || !t1) return fail(TROT_ERR_WORKSPACE, "taylor_apply_lp_symblk: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n_terms == 0) return TROT_OK;
  const int nb = (N + 7) / 8;
  const long vstride = trot_symblk_cols(N);
  HIPCHK(launch_c128_to_c64(w, t0, (long)W * N * n, s));
  float* prev = t0;
  float* next = t1;
  for (int k = 1; k <= n_terms; ++k) {
    GemmLPP<float, float, float> p{};
    p.A = vhs_symblk; p.lda = nb; p.sA = vstride;
    p.B = prev; p.ldb = n; p.sB = (long)N * n;
    p.C = next; p.ldc = n; p.sC = (long)N * n;
    p.C2 = w; p.ldc2 = n; p.sC2 = (long)N * n;
    p.M = N; p.N = n; p.K = N; p.ar = 0.f; p.ai = (float)(sqrt_dt / k);
    DISPATCH_MODE(prec, HIPCHK((launch_gemm_lp<64, 64, MODE, true, true, float, float, float, true, true, false, 1,
                                               true>(p, W, s))));
    float* tmp = prev;
    prev = next;
    next = tmp;
  }
  return TROT_OK;
}

int trot_local_energy_lp(const float* rot_chol_a, const float* rot_chol_b, const double* green_t_a,
                         const double* green_t_b, const double* rot_h1t_a, const double* rot_h1t_b, int W, int N,
                         int na, int nb, int G, double h0, double* energy, int prec, void* ws, size_t ws_bytes,
                         void* stream) {
  if (!green_t_a || !rot_h1t_a || !energy || (G > 0 && !rot_chol_a))
    return fail(TROT_ERR_INVALID, "local_energy_lp: null pointer");
  if (green_t_b && (!rot_h1t_b || (G > 0 && !rot_chol_b)))
    return fail(TROT_ERR_INVALID, "local_energy_lp: second spin incomplete");
  if (W < 1 || N < 1 || na < 1 || G < 0) return fail(TROT_ERR_INVALID, "local_energy_lp: bad sizes");
  if (na > TROT_MAX_NOCC_BIG || (green_t_b && nb > TROT_MAX_NOCC_BIG))
    return fail(TROT_ERR_UNSUPPORTED, "local_energy_lp: nocc > 256");
  if (!mode_of(prec)) return fail(TROT_ERR_INVALID, "local_energy_lp: unknown precision");
  if (ws_bytes < trot_local_energy_workspace(W, G)) return fail(TROT_ERR_WORKSPACE, "local_energy_lp: workspace");
  HIPCHK(launch_energy_lp(rot_chol_a, rot_chol_b, green_t_a, green_t_b, rot_h1t_a, rot_h1t_b, na, nb, N, G, W, h0,
                          32, (double*)ws, energy, mode_of(prec), (hipStream_t)stream));
  return TROT_OK;
}

}  // extern "C"
