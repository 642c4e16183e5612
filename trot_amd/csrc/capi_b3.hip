// C-ABI of the 3-pass split-BF16 path: operand splitting and the NT GEMM on split planes.
#include "capi_util.hpp"
#include "common.hpp"
#include "gemm_b3.hpp"

static inline long pad_to(long x, long m) { return (x + m - 1) / m * m; }

namespace {

// Grouped forms of the split, the wide GEMM's slice reduction and the split: problem i owns workgroups
// start[i] .. start[i+1] - 1 and runs the body of its own launch on its local workgroup index.
struct B3ReduceProb {
  const float* part;
  int S, M, N;
  double* C;
  long ldc;
  double ar, ai, beta;
};
struct B3SplitProb {
  const double* src;
  long ld;
  int rows, K, rp, kp;
  bf16raw* dst;
};

__device__ __forceinline__ int b3g_problem(const int* __restrict__ start, int nprob, int g) {
  int lo = 0, hi = nprob;  // start[lo] <= g < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= g) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void b3_reduce_grouped_kernel(const B3ReduceProb* __restrict__ probs,
                                                                const int* __restrict__ start, int nprob) {
  const int pi = b3g_problem(start, nprob, blockIdx.x);
  const B3ReduceProb q = probs[pi];
  b3_reduce_c128_body(q.part, q.S, q.M, q.N, q.C, q.ldc, q.ar, q.ai, q.beta,
                      (long)(blockIdx.x - start[pi]) * 256 + threadIdx.x);
}

__global__ __launch_bounds__(256) void split_c128_grouped_kernel(const B3SplitProb* __restrict__ probs,
                                                                 const int* __restrict__ start, int nprob) {
  const int pi = b3g_problem(start, nprob, blockIdx.x);
  const B3SplitProb q = probs[pi];
  split_planes_body<true, false, double>(q.src, q.ld, q.rows, q.K, q.rp, q.kp, q.dst,
                                         (long)(blockIdx.x - start[pi]) * 256 + threadIdx.x);
}

}  // namespace

extern "C" {

long trot_b3_rows_pad(int rows) { return pad_to(rows < 1 ? 1 : rows, 128); }
long trot_b3_k_pad(int K) { return pad_to(K < 1 ? 1 : K, 32); }

size_t trot_b3_planes_bytes(int rows, int K, int is_complex) {
  return (size_t)(is_complex ? 4 : 2) * trot_b3_rows_pad(rows) * trot_b3_k_pad(K) * sizeof(bf16raw);
}

// src: rows x K (row-major with leading dim ld, or column-major if transposed), FP64 or FP32
// (src_f32), real or interleaved complex.  dst: bf16 planes of (rows_pad, K_pad), zero padded.
int trot_split_bf16(const void* src, int src_f32, int is_complex, int transposed, long ld, int rows, int K,
                    void* dst, void* stream) {
  if (!src || !dst) return fail(TROT_ERR_INVALID, "split_bf16: null pointer");
  if (rows < 1 || K < 1 || ld < 1) return fail(TROT_ERR_INVALID, "split_bf16: bad sizes");
  const int rp = (int)trot_b3_rows_pad(rows), kp = (int)trot_b3_k_pad(K);
  const long total = (long)rp * kp;
  const unsigned grid = (unsigned)((total + 255) / 256);
  hipStream_t s = (hipStream_t)stream;
  bf16raw* d = (bf16raw*)dst;
#define SPLIT_LAUNCH(CX, TR, T) \
  hipLaunchKernelGGL((split_planes_kernel<CX, TR, T>), dim3(grid), dim3(256), 0, s, (const T*)src, ld, rows, K, rp, kp, d)
  if (is_complex) {
    if (src_f32) {
      if (transposed) SPLIT_LAUNCH(true, true, float); else SPLIT_LAUNCH(true, false, float);
    } else {
      if (transposed) SPLIT_LAUNCH(true, true, double); else SPLIT_LAUNCH(true, false, double);
    }
  } else {
    if (src_f32) {
      if (transposed) SPLIT_LAUNCH(false, true, float); else SPLIT_LAUNCH(false, false, float);
    } else {
      if (transposed) SPLIT_LAUNCH(false, true, double); else SPLIT_LAUNCH(false, false, double);
    }
  }
#undef SPLIT_LAUNCH
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_gemm_b3_workspace(int M, int N, int a_complex, int ksplit) {
  if (ksplit <= 1) return 0;
  return align256((size_t)ksplit * M * N * (a_complex ? 8 : 4));
}

// C (M x N, ldc; complex if a_complex, FP64 if c_f64 == 1 else FP32) = alpha * A . B^T (+ beta C);
// c_f64 == 2 (complex A, beta 0): C is split complex64, two planes of M x ldc 32-bit words
// (bf16 re | bf16 im << 16), hi = bf16(x) then lo = bf16(x - hi).
// A: split planes (M rows, K), B: split planes (N rows, K), both as made by trot_split_bf16.
// ksplit > 1 splits K over blockIdx.y with FP32 partials in ws and an FP64 reduction
// (complex FP64 output only).
int trot_gemm_b3(const void* a_planes, int a_complex, const void* b_planes, int M, int N, int K, void* C, int c_f64,
                 long ldc, double alpha_re, double alpha_im, double beta, int ksplit, void* ws, size_t ws_bytes,
                 void* stream) {
  if (!a_planes || !b_planes || !C) return fail(TROT_ERR_INVALID, "gemm_b3: null pointer");
  if (M < 1 || N < 1 || K < 1 || ldc < N) return fail(TROT_ERR_INVALID, "gemm_b3: bad sizes");
  if (!a_complex && alpha_im != 0.0) return fail(TROT_ERR_INVALID, "gemm_b3: complex alpha needs complex A");
  if (c_f64 < 0 || c_f64 > 2) return fail(TROT_ERR_INVALID, "gemm_b3: bad output kind");
  if (c_f64 == 2 && (!a_complex || beta != 0.0 || ksplit > 1))
    return fail(TROT_ERR_UNSUPPORTED, "gemm_b3: split complex64 output needs complex A, beta 0 and no split K");
  const int kp = (int)trot_b3_k_pad(K);
  if (ksplit < 1) ksplit = 1;
  int kslice = (int)pad_to((kp + ksplit - 1) / ksplit, 32);
  ksplit = (kp + kslice - 1) / kslice;
  if (ksplit > 1 && !(a_complex && c_f64))
    return fail(TROT_ERR_UNSUPPORTED, "gemm_b3: split K supports complex FP64 output only");
  if (ws_bytes < trot_gemm_b3_workspace(M, N, a_complex, ksplit))
    return fail(TROT_ERR_WORKSPACE, "gemm_b3: workspace too small");
  if (ksplit > 65535) return fail(TROT_ERR_INVALID, "gemm_b3: ksplit too large");
  B3P p{};
  p.A = (const bf16raw*)a_planes;
  p.B = (const bf16raw*)b_planes;
  p.a_plane = trot_b3_rows_pad(M) * kp;
  p.b_plane = trot_b3_rows_pad(N) * kp;
  p.Kp = kp;
  p.k_slice = kslice;
  p.M = M;
  p.N = N;
  p.C = C;
  p.ldc = ldc;
  p.part = ksplit > 1 ? (float*)ws : nullptr;
  p.ar = (float)alpha_re;
  p.ai = (float)alpha_im;
  p.beta = (float)beta;
  const int n_mt = (M + 127) / 128, n_nt = (N + 127) / 128;
  const long T = (long)n_mt * n_nt;
  if (T > 0x7fffffffL / 2) return fail(TROT_ERR_INVALID, "gemm_b3: too many tiles");
  dim3 grid((unsigned)((T + 7) / 8 * 8), (unsigned)ksplit);
  hipStream_t s = (hipStream_t)stream;
  if (a_complex) {  // wide 128 x 256 tiles
    const int n_nt2 = (N + 255) / 256;
    const long T2 = (long)n_mt * n_nt2;
    dim3 grid2((unsigned)((T2 + 7) / 8 * 8), (unsigned)ksplit);
    const int b_rows = (int)trot_b3_rows_pad(N);
    if (c_f64 == 2)
      hipLaunchKernelGGL((gemm_b3w_kernel<unsigned>), grid2, dim3(512), 0, s, p, n_mt, n_nt2, b_rows);
    else if (c_f64)
      hipLaunchKernelGGL((gemm_b3w_kernel<double>), grid2, dim3(512), 0, s, p, n_mt, n_nt2, b_rows);
    else
      hipLaunchKernelGGL((gemm_b3w_kernel<float>), grid2, dim3(512), 0, s, p, n_mt, n_nt2, b_rows);
  } else {
    if (c_f64)
      hipLaunchKernelGGL((gemm_b3_nt_kernel<false, double>), grid, dim3(256), 0, s, p, n_mt, n_nt);
    else
      hipLaunchKernelGGL((gemm_b3_nt_kernel<false, float>), grid, dim3(256), 0, s, p, n_mt, n_nt);
  }
  HIPCHK(hipGetLastError());
  if (ksplit > 1) {
    const long total = (long)M * N;
    hipLaunchKernelGGL(b3_reduce_c128_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       (const float*)ws, ksplit, M, N, (double*)C, ldc, alpha_re, alpha_im, beta);
    HIPCHK(hipGetLastError());
  }
  return TROT_OK;
}


// ---- grouped forms for many small problems in one launch (the batching form in mixed precision) ----

size_t trot_split_bf16_problem_bytes(void) { return sizeof(B3SplitProb); }

// One complex FP64 row-major operand (rows x K, leading dim ld) to split as trot_split_bf16 does.  Returns the
// problem's workgroup count.
long trot_split_bf16_problem(const void* src, long ld, int rows, int K, void* dst, void* desc) {
  if (!src || !dst || !desc) return fail(TROT_ERR_INVALID, "split_bf16_problem: null pointer");
  if (rows < 1 || K < 1 || ld < K) return fail(TROT_ERR_INVALID, "split_bf16_problem: bad sizes");
  B3SplitProb q{};
  q.src = (const double*)src;
  q.ld = ld;
  q.rows = rows;
  q.K = K;
  q.rp = (int)trot_b3_rows_pad(rows);
  q.kp = (int)trot_b3_k_pad(K);
  q.dst = (bf16raw*)dst;
  *reinterpret_cast<B3SplitProb*>(desc) = q;
  return ((long)q.rp * q.kp + 255) / 256;
}

int trot_split_bf16_grouped(const void* probs, const int* start, int nprob, int total_wg, void* stream) {
  if (!probs || !start || nprob < 1 || total_wg < 1) return fail(TROT_ERR_INVALID, "split_bf16_grouped: bad arguments");
  hipLaunchKernelGGL(split_c128_grouped_kernel, dim3(total_wg), dim3(256), 0, (hipStream_t)stream,
                     (const B3SplitProb*)probs, start, nprob);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_gemm_b3_problem_bytes(void) { return sizeof(B3GProb); }

// One product of trot_gemm_b3 with complex A (the wide 128 x 256 kernel) as a grouped problem: the same
// arguments and the same tiles, K slices and partial sums (in ws).  Returns the workgroup count.
long trot_gemm_b3_problem(const void* a_planes, const void* b_planes, int M, int N, int K, void* C, int c_kind,
                          long ldc, double alpha_re, double alpha_im, double beta, int ksplit, void* ws,
                          size_t ws_bytes, void* desc) {
  if (!a_planes || !b_planes || !C || !desc) return fail(TROT_ERR_INVALID, "gemm_b3_problem: null pointer");
  if (M < 1 || N < 1 || K < 1 || ldc < N) return fail(TROT_ERR_INVALID, "gemm_b3_problem: bad sizes");
  if (c_kind < 0 || c_kind > 2) return fail(TROT_ERR_INVALID, "gemm_b3_problem: bad output kind");
  if (c_kind == 2 && (beta != 0.0 || ksplit > 1))
    return fail(TROT_ERR_UNSUPPORTED, "gemm_b3_problem: split complex64 output needs beta 0 and no split K");
  const int kp = (int)trot_b3_k_pad(K);
  if (ksplit < 1) ksplit = 1;
  int kslice = (int)pad_to((kp + ksplit - 1) / ksplit, 32);
  ksplit = (kp + kslice - 1) / kslice;
  if (ksplit > 1 && c_kind != 1) return fail(TROT_ERR_UNSUPPORTED, "gemm_b3_problem: split K needs FP64 output");
  if (ws_bytes < trot_gemm_b3_workspace(M, N, 1, ksplit)) return fail(TROT_ERR_WORKSPACE, "gemm_b3_problem: workspace");
  B3GProb q{};
  B3P& p = q.p;
  p.A = (const bf16raw*)a_planes;
  p.B = (const bf16raw*)b_planes;
  p.a_plane = trot_b3_rows_pad(M) * kp;
  p.b_plane = trot_b3_rows_pad(N) * kp;
  p.Kp = kp;
  p.k_slice = kslice;
  p.M = M;
  p.N = N;
  p.C = C;
  p.ldc = ldc;
  p.part = ksplit > 1 ? (float*)ws : nullptr;
  p.ar = (float)alpha_re;
  p.ai = (float)alpha_im;
  p.beta = (float)beta;
  q.n_mt = (M + 127) / 128;
  q.n_nt = (N + 255) / 256;
  q.b_rows = (int)trot_b3_rows_pad(N);
  q.S = ksplit;
  const long T2 = (long)q.n_mt * q.n_nt;
  if (T2 * ksplit > 0x7fffffffL / 2) return fail(TROT_ERR_INVALID, "gemm_b3_problem: too many tiles");
  q.tiles_pad = (int)((T2 + 7) / 8 * 8);
  q.c_kind = c_kind;
  q.ar = alpha_re;
  q.ai = alpha_im;
  q.beta = beta;
  *reinterpret_cast<B3GProb*>(desc) = q;
  return (long)q.tiles_pad * ksplit;
}

// The wide GEMMs of nprob problems of one output kind (c_kind) in one launch.
int trot_gemm_b3_grouped(int c_kind, const void* probs, const int* start, int nprob, int total_wg, void* stream) {
  if (!probs || !start || nprob < 1 || total_wg < 1) return fail(TROT_ERR_INVALID, "gemm_b3_grouped: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const B3GProb* q = (const B3GProb*)probs;
  if (c_kind == 2)
    hipLaunchKernelGGL((gemm_b3w_grouped_kernel<unsigned>), dim3(total_wg), dim3(512), 0, s, q, start, nprob);
  else if (c_kind == 1)
    hipLaunchKernelGGL((gemm_b3w_grouped_kernel<double>), dim3(total_wg), dim3(512), 0, s, q, start, nprob);
  else if (c_kind == 0)
    hipLaunchKernelGGL((gemm_b3w_grouped_kernel<float>), dim3(total_wg), dim3(512), 0, s, q, start, nprob);
  else
    return fail(TROT_ERR_INVALID, "gemm_b3_grouped: bad output kind");
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

size_t trot_b3_reduce_problem_bytes(void) { return sizeof(B3ReduceProb); }

// The slice reduction of a split-K problem from trot_gemm_b3_problem (C = alpha sum of slices + beta C, in
// slice order, FP64), as a grouped problem.  Returns its workgroup count, 0 for a product without split K.
long trot_b3_reduce_problem(const void* gemm_desc, void* desc) {
  if (!gemm_desc || !desc) return fail(TROT_ERR_INVALID, "b3_reduce_problem: null pointer");
  const B3GProb& g = *reinterpret_cast<const B3GProb*>(gemm_desc);
  if (g.S <= 1) return 0;
  B3ReduceProb r{};
  r.part = g.p.part;
  r.S = g.S;
  r.M = g.p.M;
  r.N = g.p.N;
  r.C = (double*)g.p.C;
  r.ldc = g.p.ldc;
  r.ar = g.ar;
  r.ai = g.ai;
  r.beta = g.beta;
  *reinterpret_cast<B3ReduceProb*>(desc) = r;
  return ((long)r.M * r.N + 255) / 256;
}

int trot_b3_reduce_grouped(const void* probs, const int* start, int nprob, int total_wg, void* stream) {
  if (!probs || !start || nprob < 1 || total_wg < 1) return fail(TROT_ERR_INVALID, "b3_reduce_grouped: bad arguments");
  hipLaunchKernelGGL(b3_reduce_grouped_kernel, dim3(total_wg), dim3(256), 0, (hipStream_t)stream,
                     (const B3ReduceProb*)probs, start, nprob);
  HIPCHK(hipGetLastError());
  return TROT_OK;
}

}  // extern "C"
