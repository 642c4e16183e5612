// General batched FP64 GEMM and batched inverse entry points: the matrix products of the hand-written adjoint
// of the phaseless step (Green's-function, Taylor-series and one-body adjoints), on the FP64 MFMA tile of
// gemm.hpp.
//
//   C[b] = alpha op(A[b]) op(B[b]) + beta C[b],   op in {N, T, C = conjugate transpose, R = conjugate}
//
// A is M x K after op, B is K x N after op; a complex operand is interleaved (re, im) doubles; the product is
// complex whenever either operand is.  Strides and leading dimensions count elements (complex elements for
// complex operands); a batch stride of 0 shares the operand across the batch.
#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

#include "capi_util.hpp"

namespace {

// Storage of op(X) for the tile: K-contiguous A / N-contiguous B for N and R, the other way for T and C.
// Products with at most 32 output columns (the walker side of n <= 32) run on 64 x 32 tiles, which do half
// the MFMA work of a 64 x 64 tile on them; every tile shape accumulates K in the same order.
template <bool AC, bool BC, bool AKC, bool BNC, bool CJA, bool CJB>
hipError_t run(const GemmP& p, int batch, hipStream_t s) {
  if (p.N <= 32) return launch_gemm_f64<64, 32, AC, BC, AKC, BNC, CJA, EPI_STORE, false, CJB>(p, batch, s);
  return launch_gemm_f64<64, 64, AC, BC, AKC, BNC, CJA, EPI_STORE, false, CJB>(p, batch, s);
}

template <bool AC, bool BC, bool AKC, bool CJA>
hipError_t pick_b(int opb, const GemmP& p, int batch, hipStream_t s) {
  if (!BC) return opb == TROT_OP_N ? run<AC, false, AKC, true, CJA, false>(p, batch, s)
                                   : run<AC, false, AKC, false, CJA, false>(p, batch, s);
  switch (opb) {
    case TROT_OP_N: return run<AC, BC, AKC, true, CJA, false>(p, batch, s);
    case TROT_OP_T: return run<AC, BC, AKC, false, CJA, false>(p, batch, s);
    case TROT_OP_C: return run<AC, BC, AKC, false, CJA, true>(p, batch, s);
    default: return run<AC, BC, AKC, true, CJA, true>(p, batch, s);  // TROT_OP_R
  }
}

template <bool AC, bool BC>
hipError_t pick_a(int opa, int opb, const GemmP& p, int batch, hipStream_t s) {
  if (!AC) return opa == TROT_OP_N ? pick_b<false, BC, true, false>(opb, p, batch, s)
                                   : pick_b<false, BC, false, false>(opb, p, batch, s);
  switch (opa) {
    case TROT_OP_N: return pick_b<AC, BC, true, false>(opb, p, batch, s);
    case TROT_OP_T: return pick_b<AC, BC, false, false>(opb, p, batch, s);
    case TROT_OP_C: return pick_b<AC, BC, false, true>(opb, p, batch, s);
    default: return pick_b<AC, BC, true, true>(opb, p, batch, s);  // TROT_OP_R
  }
}

}  // namespace

extern "C" {

int trot_gemm_batched(int opa, int opb, int a_complex, int b_complex, int M, int N, int K, double alpha_re,
                      double alpha_im, const double* A, long lda, long sA, const double* B, long ldb, long sB,
                      double beta, double* C, long ldc, long sC, int batch, void* stream) {
  if (!A || !B || !C) return fail(TROT_ERR_INVALID, "gemm_batched: null pointer");
  if (opa < TROT_OP_N || opa > TROT_OP_R || opb < TROT_OP_N || opb > TROT_OP_R)
    return fail(TROT_ERR_INVALID, "gemm_batched: unknown op");
  if ((!a_complex && opa >= TROT_OP_C) || (!b_complex && opb >= TROT_OP_C))
    return fail(TROT_ERR_INVALID, "gemm_batched: conjugation of a real operand");
  if (M < 0 || N < 0 || K < 0 || batch < 0 || sA < 0 || sB < 0 || sC < 0)
    return fail(TROT_ERR_INVALID, "gemm_batched: bad sizes");
  const bool a_kc = opa == TROT_OP_N || opa == TROT_OP_R;  // A stored M x K
  const bool b_nc = opb == TROT_OP_N || opb == TROT_OP_R;  // B stored K x N
  if (lda < (a_kc ? K : M) || ldb < (b_nc ? N : K) || ldc < N) return fail(TROT_ERR_INVALID, "gemm_batched: leading dimension");
  // the FP64 tile GEMM runs one workgroup per (batch, 64-row tile, 32- or 64-column tile) on a 1D grid
  if ((long)((M + 63) / 64) * ((N + 31) / 32) * batch > 0x7ffffff0L)
    return fail(TROT_ERR_UNSUPPORTED, "gemm_batched: too many tiles");
  if (!(a_complex || b_complex) && alpha_im != 0.0)
    return fail(TROT_ERR_INVALID, "gemm_batched: complex alpha with a real product");
  if (M == 0 || N == 0 || batch == 0) return TROT_OK;
  GemmP p{};
  p.A = A; p.lda = lda; p.sA = sA;
  p.B = B; p.ldb = ldb; p.sB = sB;
  p.C = C; p.ldc = ldc; p.sC = sC;
  p.M = M; p.N = N; p.K = K; p.ar = alpha_re; p.ai = alpha_im; p.beta = beta;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (a_complex && b_complex) e = pick_a<true, true>(opa, opb, p, batch, s);
  else if (a_complex) e = pick_a<true, false>(opa, opb, p, batch, s);
  else if (b_complex) e = pick_a<false, true>(opa, opb, p, batch, s);
  else e = pick_a<false, false>(opa, opb, p, batch, s);
  HIPCHK(e);
  return TROT_OK;
}

int trot_inverse_det(const double* m, int W, int n, double* minv, double* det, void* stream) {
  if (!m || !minv || !det || m == minv) return fail(TROT_ERR_INVALID, "inverse_det: bad pointers");
  if (W < 0 || n < 1) return fail(TROT_ERR_INVALID, "inverse_det: bad sizes");
  if (n > TROT_MAX_NOCC_BIG) return fail(TROT_ERR_UNSUPPORTED, "inverse_det: n > 256");
  HIPCHK(launch_inv_det(m, minv, det, n, W, (hipStream_t)stream));
  return TROT_OK;
}

}  // extern "C"
