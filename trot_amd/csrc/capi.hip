// C-ABI of trot_amd: extern "C" entry points over plain device pointers.
//
// Conventions (also in include/trot_amd.h):
//  * complex arrays are interleaved (re, im) doubles (complex128 layout);
//  * every entry point returns TROT_OK (0) or a negative status and never
//    throws; trot_last_error() gives a message for the calling thread;
//  * the caller owns every buffer; scratch memory comes in through a
//    (ws, ws_bytes) pair sized by the matching *_workspace() query.  No entry
//    point allocates device memory;
//  * `stream` is a hipStream_t (NULL = default stream); all work is enqueued
//    on it asynchronously.
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

#include "common.hpp"
#include "gemm.hpp"
#include "kernels.hpp"

#include "capi_util.hpp"

thread_local char g_trot_err[512];

// ---------------------------------------------------------------------------
// Static scratch of the allocation-free comb (trot_sr_comb_dev_noalloc): the cumulative weights of up to
// TROT_SR_NOALLOC_MAX_W walkers.  Each GPU holds its own copy of the array until the process ends.  In host
// code `scratch` is only a handle: its device address on the current GPU comes from hipGetSymbolAddress,
// looked up once per device and kept in g_scratch_dev.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[TROT_SR_NOALLOC_MAX_W];
static std::mutex g_scratch_mu;
static std::map<int, double*> g_scratch_dev;  // device ordinal -> address of `scratch` on that device

static int scratch_on_current_device(double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_dev.find(dev);
  if (it == g_scratch_dev.end()) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TROT_ERR_HIP, "sr_comb_dev_noalloc: no device address for the static scratch");
    it = g_scratch_dev.emplace(dev, static_cast<double*>(p)).first;
  }
  *out = it->second;
  return TROT_OK;
}

extern "C" {

const char* trot_version(void) { return "trot_amd 0.1 (gfx950, FP64 MFMA)"; }
const char* trot_last_error(void) { return g_trot_err; }

// ---------------------------------------------------------------------------
// overlap + half Green's function
// ---------------------------------------------------------------------------
size_t trot_overlap_green_workspace(int W, int N, int n) {
  (void)N;
  return 2 * align256((size_t)W * n * n * 16);
}

// ---------------------------------------------------------------------------
// GEMM problems of the step, built in one place so that a grouped launch over many walker sets runs
// exactly the problem a launch of its own would (same operands, strides and tile order of K).
// ---------------------------------------------------------------------------
static GemmP problem_overlap(const double* trial, const double* walkers, double* m, int N, int n) {
  GemmP p{};  // m[b] = C^T w[b]  (n x n)
  p.A = trial; p.lda = n; p.sA = 0;
  p.B = walkers; p.ldb = n; p.sB = (long)N * n;
  p.C = m; p.ldc = n; p.sC = (long)n * n;
  p.M = n; p.N = n; p.K = N; p.ar = 1.0;
  return p;
}
static GemmP problem_green(const double* walkers, const double* minv, double* green_t, int N, int n) {
  GemmP q{};  // GT[b] = w[b] m[b]^-1  (N x n)
  q.A = walkers; q.lda = n; q.sA = (long)N * n;
  q.B = minv; q.ldb = n; q.sB = (long)n * n;
  q.C = green_t; q.ldc = n; q.sC = (long)N * n;
  q.M = N; q.N = n; q.K = n; q.ar = 1.0;
  return q;
}
static GemmP problem_vhs(const double* xbar, const double* chol, double* vhs, int G, long N2, int W) {
  GemmP p{};
  p.A = xbar; p.lda = G;
  p.B = chol; p.ldb = N2;
  p.C = vhs; p.ldc = N2;
  p.M = W; p.N = (int)N2; p.K = G; p.ar = 1.0;
  return p;
}
static GemmP problem_one_body(const double* expm, const double* w_in, double* w_out, int N, int n) {
  GemmP p{};
  p.A = expm; p.lda = N; p.sA = 0;
  p.B = w_in; p.ldb = n; p.sB = (long)N * n;
  p.C = w_out; p.ldc = n; p.sC = (long)N * n;
  p.M = N; p.N = n; p.K = N; p.ar = 1.0;
  return p;
}
static GemmP problem_taylor_symblk(const double* vhs_symblk, const double* prev, double* next, double* acc, int N,
                                   int n, double coef) {
  GemmP p{};
  p.A = vhs_symblk; p.lda = (N + 7) / 8; p.sA = symblk_count((N + 7) / 8) * 64;
  p.B = prev; p.ldb = n; p.sB = (long)N * n;
  p.C = next; p.ldc = n; p.sC = (long)N * n;
  p.C2 = acc; p.ldc2 = n; p.sC2 = (long)N * n;
  p.M = N; p.N = n; p.K = N; p.ar = 0.0; p.ai = coef;
  return p;
}

int trot_overlap_green(const double* trial, const double* walkers, int W, int N, int n, double* det,
                       double* green_t, void* ws, size_t ws_bytes, void* stream) {
  if (!trial || !walkers || !det) return fail(TROT_ERR_INVALID, "overlap_green: null pointer");
  if (W < 1 || N < 1 || n < 1 || n > N) return fail(TROT_ERR_INVALID, "overlap_green: bad sizes");
  if (n > TROT_MAX_NOCC_BIG) return fail(TROT_ERR_UNSUPPORTED, "overlap_green: nocc > 256 unsupported");
  Carve c{(char*)ws, ws_bytes, 0};
  double* m = c.take<double>((size_t)W * n * n * 2);
  double* minv = c.take<double>((size_t)W * n * n * 2);
  if (!m || !minv) return fail(TROT_ERR_WORKSPACE, "overlap_green: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  HIPCHK((launch_gemm_f64<64, 64, false, true, false, true, false, EPI_STORE>(
      problem_overlap(trial, walkers, m, N, n), W, s)));
  HIPCHK(launch_inv_det(m, minv, det, n, W, s));
  if (green_t)
    HIPCHK((launch_gemm_f64<64, 64, true, true, true, true, false, EPI_STORE>(
        problem_green(walkers, minv, green_t, N, n), W, s)));
  return TROT_OK;
}

int trot_overlap_combine(const double* det_a, const double* det_b, double* overlap, int W, void* stream) {
  if (!det_a || !overlap || W < 0) return fail(TROT_ERR_INVALID, "overlap_combine: bad arguments");
  HIPCHK(launch_overlap_combine(det_a, det_b, overlap, W, (hipStream_t)stream));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// force bias: fb (+)= alpha * GT_flat (W x Nn) . R'^T, R' = (G, Nn) real
// ---------------------------------------------------------------------------
int trot_force_bias(const double* green_t, const double* rot_chol_t, int W, int G, long Nn, double alpha,
                    int accumulate, double* fb, void* stream) {
  if (!green_t || !rot_chol_t || !fb) return fail(TROT_ERR_INVALID, "force_bias: null pointer");
  if (W < 1 || G < 1 || Nn < 1) return fail(TROT_ERR_INVALID, "force_bias: bad sizes");
  GemmP p{};
  p.A = green_t; p.lda = Nn; p.sA = 0;
  p.B = rot_chol_t; p.ldb = Nn; p.sB = 0;
  p.C = fb; p.ldc = G; p.sC = 0;
  p.M = W; p.N = G; p.K = (int)Nn; p.ar = alpha; p.beta = accumulate ? 1.0 : 0.0;
  HIPCHK((launch_gemm_f64<64, 64, true, false, true, false, false, EPI_STORE>(p, 1, (hipStream_t)stream)));
  return TROT_OK;
}

// Split K: S slices of K, FP64 partial products in ws, summed in slice order.  S is a function of the
// shape only (same shape, same S, same rounding): enough slices for about 2048 workgroups (8 per CU; at
// W = 2048, G = 2560, N n = 32768 two slices of the 1280 tiles run the FP64 force bias 3.4 % faster per step
// than none, and a 1024 target left it unsplit), each slice at least 256 long, at most 64 slices.
static int force_bias_slices(int W, int G, long Nn) {
  const long tiles = (long)((G + 63) / 64) * ((W + 63) / 64);
  long S = (2048 + tiles - 1) / tiles;
  const long smax = Nn / 256;
  if (S > smax) S = smax;
  if (S > 64) S = 64;
  return S < 1 ? 1 : (int)S;
}

size_t trot_force_bias_workspace(int W, int G, long Nn) {
  if (W < 1 || G < 1 || Nn < 1) return 0;
  const int S = force_bias_slices(W, G, Nn);
  return S > 1 ? align256((size_t)S * W * G * 16) : 0;
}

int trot_force_bias_splitk(const double* green_t, const double* rot_chol_t, int W, int G, long Nn, double alpha,
                           int accumulate, double* fb, void* ws, size_t ws_bytes, void* stream) {
  if (!green_t || !rot_chol_t || !fb) return fail(TROT_ERR_INVALID, "force_bias: null pointer");
  if (W < 1 || G < 1 || Nn < 1) return fail(TROT_ERR_INVALID, "force_bias: bad sizes");
  const int S = force_bias_slices(W, G, Nn);
  if (S == 1) return trot_force_bias(green_t, rot_chol_t, W, G, Nn, alpha, accumulate, fb, stream);
  if (!ws || ws_bytes < trot_force_bias_workspace(W, G, Nn)) return fail(TROT_ERR_WORKSPACE, "force_bias: workspace");
  const long ks = (Nn + S - 1) / S;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)ws;
  GemmP p{};
  p.A = green_t; p.lda = Nn; p.sA = ks;
  p.B = rot_chol_t; p.ldb = Nn; p.sB = ks;
  p.C = part; p.ldc = G; p.sC = (long)W * G;
  p.M = W; p.N = G; p.K = (int)ks; p.Klast = (int)(Nn - (S - 1) * ks); p.ar = 1.0;
  HIPCHK((launch_gemm_f64<64, 64, true, false, true, false, false, EPI_STORE>(p, S, s)));
  HIPCHK(launch_sum_batches(part, fb, S, 2L * W * G, alpha, accumulate, s));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// shifted fields
// ---------------------------------------------------------------------------
int trot_field_shift(const double* fields, const double* fb, const double* mf, double sqrt_dt, int W, int G,
                     double* xbar, double* shift_term, double* fb_term, void* stream) {
  if (!fields || !fb || !mf || !xbar || !shift_term || !fb_term)
    return fail(TROT_ERR_INVALID, "field_shift: null pointer");
  if (W < 1 || G < 1) return fail(TROT_ERR_INVALID, "field_shift: bad sizes");
  HIPCHK(launch_shift(fields, fb, mf, sqrt_dt, W, G, xbar, nullptr, shift_term, fb_term, (hipStream_t)stream));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// VHS: V (W x N^2 complex) = xbar (W x G complex) . Lflat (G x N^2 real)
// ---------------------------------------------------------------------------
int trot_vhs_build(const double* xbar, const double* chol_flat, int W, int G, long N2, double* vhs, void* stream) {
  if (!xbar || !chol_flat || !vhs) return fail(TROT_ERR_INVALID, "vhs_build: null pointer");
  if (W < 1 || G < 1 || N2 < 1) return fail(TROT_ERR_INVALID, "vhs_build: bad sizes");
  HIPCHK((launch_gemm_f64<64, 64, true, false, true, true, false, EPI_STORE>(problem_vhs(xbar, chol_flat, vhs, G, N2, W),
                                                                            1, (hipStream_t)stream)));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// one-body half step: w_out[b] = E (N x N real) . w_in[b]
// ---------------------------------------------------------------------------
int trot_one_body_apply(const double* expm, const double* w_in, int W, int N, int n, double* w_out, void* stream) {
  if (!expm || !w_in || !w_out) return fail(TROT_ERR_INVALID, "one_body_apply: null pointer");
  if (w_in == w_out) return fail(TROT_ERR_INVALID, "one_body_apply: in-place not supported");
  if (W < 1 || N < 1 || n < 1) return fail(TROT_ERR_INVALID, "one_body_apply: bad sizes");
  HIPCHK((launch_gemm_f64<64, 64, false, true, true, true, false, EPI_STORE>(problem_one_body(expm, w_in, w_out, N, n),
                                                                             W, (hipStream_t)stream)));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// Taylor action: w <- sum_{k=0..K} (i sqrt(dt) V)^k / k! w  (term recursion)
// ---------------------------------------------------------------------------
size_t trot_taylor_workspace(int W, int N, int n) { return 2 * align256((size_t)W * N * n * 16); }

// Tile shape of the FP64 Taylor-term GEMM (M = N orbitals, N = n occupied, one product per walker).
// Every shape accumulates K in the same ascending order of 4-wide MFMA steps, so the choice changes
// speed only, never a bit of the result.  Choices come from trot_autotune_taylor (timed on the caller's
// stream) or trot_set_taylor_tile and live in one process-wide table behind a mutex; the default is 64x64.
enum { TILE_64x64 = 0, TILE_64x32 = 1, TILE_32x64 = 2, TILE_32x32 = 3, TILE_COUNT = 4 };
static std::mutex g_tile_mu;
static std::map<std::tuple<int, int, int, int>, int> g_tile;  // (storage, W, N, n) -> shape

static int taylor_tile(int sym, int W, int N, int n) {
  std::lock_guard<std::mutex> lk(g_tile_mu);
  auto it = g_tile.find(std::make_tuple(sym, W, N, n));
  return it == g_tile.end() ? TILE_64x64 : it->second;
}

extern "C++" {
template <bool ASYM>
static hipError_t launch_taylor_term(int shape, const GemmP& p, int W, hipStream_t s) {
  switch (shape) {
    case TILE_64x32: return launch_gemm_f64<64, 32, true, true, true, true, false, EPI_TAYLOR, ASYM>(p, W, s);
    case TILE_32x64: return launch_gemm_f64<32, 64, true, true, true, true, false, EPI_TAYLOR, ASYM>(p, W, s);
    case TILE_32x32: return launch_gemm_f64<32, 32, true, true, true, true, false, EPI_TAYLOR, ASYM>(p, W, s);
    default: return launch_gemm_f64<64, 64, true, true, true, true, false, EPI_TAYLOR, ASYM>(p, W, s);
  }
}
}  // extern "C++"

int trot_taylor_apply(const double* vhs, double* w, int W, int N, int n, double sqrt_dt, int n_terms, void* ws,
                      size_t ws_bytes, void* stream) {
  if (!vhs || !w) return fail(TROT_ERR_INVALID, "taylor_apply: null pointer");
  if (W < 1 || N < 1 || n < 1 || n_terms < 0) return fail(TROT_ERR_INVALID, "taylor_apply: bad sizes");
  Carve c{(char*)ws, ws_bytes, 0};
  double* t0 = c.take<double>((size_t)W * N * n * 2);
  double* t1 = c.take<double>((size_t)W * N * n * 2);
  if (!t0 || !t1) return fail(TROT_ERR_WORKSPACE, "taylor_apply: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n_terms == 0) return TROT_OK;
  const int shape = taylor_tile(0, W, N, n);
  HIPCHK(hipMemcpyAsync(t0, w, (size_t)W * N * n * 16, hipMemcpyDeviceToDevice, s));
  double* prev = t0;
  double* next = t1;
  for (int k = 1; k <= n_terms; ++k) {
    GemmP p{};
    p.A = vhs; p.lda = N; p.sA = (long)N * N;
    p.B = prev; p.ldb = n; p.sB = (long)N * n;
    p.C = next; p.ldc = n; p.sC = (long)N * n;
    p.C2 = w; p.ldc2 = n; p.sC2 = (long)N * n;
    p.M = N; p.N = n; p.K = N; p.ar = 0.0; p.ai = sqrt_dt / k;
    HIPCHK(launch_taylor_term<false>(shape, p, W, s));
    double* tmp = prev;
    prev = next;
    next = tmp;
  }
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// symmetric Cholesky vectors: sym-block storage of L_g and of V
// ---------------------------------------------------------------------------
long trot_symblk_cols(int N) {
  const int nb = (N + 7) / 8;
  return symblk_count(nb) * 64;
}

int trot_pack_symblk(const double* chol, int G, int N, void* out, int out_f32, void* stream) {
  if (!chol || !out) return fail(TROT_ERR_INVALID, "pack_symblk: null pointer");
  if (G < 1 || N < 1 || G > 65535) return fail(TROT_ERR_INVALID, "pack_symblk: bad sizes");
  HIPCHK(launch_pack_symblk(chol, out, out_f32, G, N, (hipStream_t)stream));
  return TROT_OK;
}

int trot_taylor_apply_symblk(const double* vhs_symblk, double* w, int W, int N, int n, double sqrt_dt, int n_terms,
                             void* ws, size_t ws_bytes, void* stream) {
  if (!vhs_symblk || !w) return fail(TROT_ERR_INVALID, "taylor_apply_symblk: null pointer");
  if (W < 1 || N < 1 || n < 1 || n_terms < 0) return fail(TROT_ERR_INVALID, "taylor_apply_symblk: bad sizes");
  Carve c{(char*)ws, ws_bytes, 0};
  double* t0 = c.take<double>((size_t)W * N * n * 2);
  double* t1 = c.take<double>((size_t)W * N * n * 2);
  if (!t0 || !t1) return fail(TROT_ERR_WORKSPACE, "taylor_apply_symblk: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n_terms == 0) return TROT_OK;
  const int shape = taylor_tile(1, W, N, n);
  HIPCHK(hipMemcpyAsync(t0, w, (size_t)W * N * n * 16, hipMemcpyDeviceToDevice, s));
  const int nb = (N + 7) / 8;
  double* prev = t0;
  double* next = t1;
  (void)nb;
  for (int k = 1; k <= n_terms; ++k) {
    HIPCHK(launch_taylor_term<true>(shape, problem_taylor_symblk(vhs_symblk, prev, next, w, N, n, sqrt_dt / k), W, s));
    double* tmp = prev;
    prev = next;
    next = tmp;
  }
  return TROT_OK;
}

size_t trot_autotune_taylor_workspace(int W, int N, int n, int symblk) {
  if (W < 1 || N < 1 || n < 1) return 0;
  const size_t v = (size_t)W * (symblk ? trot_symblk_cols(N) : (long)N * N) * 16;
  return align256(v) + 3 * align256((size_t)W * N * n * 16);
}

int trot_set_taylor_tile(int W, int N, int n, int symblk, int shape) {
  if (shape < 0 || shape >= TILE_COUNT) return fail(TROT_ERR_INVALID, "set_taylor_tile: unknown shape");
  std::lock_guard<std::mutex> lk(g_tile_mu);
  g_tile[std::make_tuple(symblk ? 1 : 0, W, N, n)] = shape;
  return TROT_OK;
}

int trot_get_taylor_tile(int W, int N, int n, int symblk) { return taylor_tile(symblk ? 1 : 0, W, N, n); }

// Times one Taylor term (K = 1: one product and the accumulation) in each tile shape on scratch
// operands carved from ws (filled with values in [-1, 1)), three launches per shape after one warm-up, and records the fastest for
// (W, N, n, storage).  Synchronizes the stream: not for use while the stream is being captured.
int trot_autotune_taylor(int W, int N, int n, int symblk, void* ws, size_t ws_bytes, void* stream, int* shape_out,
                         float* ms_out) {
  if (W < 1 || N < 1 || n < 1) return fail(TROT_ERR_INVALID, "autotune_taylor: bad sizes");
  if (ws_bytes < trot_autotune_taylor_workspace(W, N, n, symblk))
    return fail(TROT_ERR_WORKSPACE, "autotune_taylor: workspace");
  hipStream_t s = (hipStream_t)stream;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  HIPCHK(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return fail(TROT_ERR_INVALID, "autotune_taylor: stream is being captured");
  Carve c{(char*)ws, ws_bytes, 0};
  const long vcols = symblk ? trot_symblk_cols(N) : (long)N * N;
  double* v = c.take<double>((size_t)W * vcols * 2);
  double* w = c.take<double>((size_t)W * N * n * 2);
  double* t0 = c.take<double>((size_t)W * N * n * 2);
  double* t1 = c.take<double>((size_t)W * N * n * 2);
  // operands of size O(1) (a hashed pattern): MFMA timings on all-zero operands run fast and mislead
  HIPCHK(launch_fill_pattern((double*)ws, (long)(trot_autotune_taylor_workspace(W, N, n, symblk) / 8), s));
  GemmP p{};
  p.A = v; p.lda = symblk ? (N + 7) / 8 : N; p.sA = vcols;
  p.B = t0; p.ldb = n; p.sB = (long)N * n;
  p.C = t1; p.ldc = n; p.sC = (long)N * n;
  p.C2 = w; p.ldc2 = n; p.sC2 = (long)N * n;
  p.M = N; p.N = n; p.K = N; p.ar = 0.0; p.ai = 0.1;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int best = TILE_64x64;
  float best_ms = 0.f;
  int rc = TROT_OK;
  for (int shape = 0; shape < TILE_COUNT && rc == TROT_OK; ++shape) {
    hipError_t e = symblk ? launch_taylor_term<true>(shape, p, W, s) : launch_taylor_term<false>(shape, p, W, s);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < 3 && e == hipSuccess; ++r)
      e = symblk ? launch_taylor_term<true>(shape, p, W, s) : launch_taylor_term<false>(shape, p, W, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) {
      rc = hip_status(e, "autotune_taylor");
      break;
    }
    ms /= 3.f;
    if (ms_out) ms_out[shape] = ms;
    if (shape == 0 || ms < best_ms) {
      best = shape;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != TROT_OK) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tile_mu);
    g_tile[std::make_tuple(symblk ? 1 : 0, W, N, n)] = best;
  }
  if (shape_out) *shape_out = best;
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// full Trotter factor for spin sets sharing one V, in walker chunks:
//   w_s <- E . T_K(i sqrt(dt) V(xbar)) . E . w_s
// walkers_b may be NULL (restricted walkers).
// ---------------------------------------------------------------------------
size_t trot_apply_trotter_workspace(int W, int N, int n_max, int chunk) {
  if (chunk <= 0 || chunk > W) chunk = W;
  const size_t v = align256((size_t)chunk * N * N * 16);
  const size_t wb = align256((size_t)chunk * N * n_max * 16);
  return v + wb + trot_taylor_workspace(chunk, N, n_max);
}

int trot_apply_trotter(double* walkers_a, int na, double* walkers_b, int nb, const double* xbar,
                       const double* chol_flat, const double* expm, int W, int N, int G, double sqrt_dt,
                       int n_terms, int chunk, void* ws, size_t ws_bytes, void* stream) {
  if (!walkers_a || !xbar || !chol_flat || !expm) return fail(TROT_ERR_INVALID, "apply_trotter: null pointer");
  if (W < 1 || N < 1 || G < 1 || na < 1 || (walkers_b && nb < 1))
    return fail(TROT_ERR_INVALID, "apply_trotter: bad sizes");
  if (chunk <= 0 || chunk > W) chunk = W;
  const int n_max = (walkers_b && nb > na) ? nb : na;
  if (ws_bytes < trot_apply_trotter_workspace(W, N, n_max, chunk))
    return fail(TROT_ERR_WORKSPACE, "apply_trotter: workspace too small");
  Carve c{(char*)ws, ws_bytes, 0};
  double* v = c.take<double>((size_t)chunk * N * N * 2);
  double* wtmp = c.take<double>((size_t)chunk * N * n_max * 2);
  void* tws = c.base + c.used;
  const size_t tws_bytes = c.cap - c.used;
  hipStream_t s = (hipStream_t)stream;
  (void)s;
  for (int w0 = 0; w0 < W; w0 += chunk) {
    const int wc = (W - w0) < chunk ? (W - w0) : chunk;
    int rc = trot_vhs_build(xbar + 2 * (size_t)w0 * G, chol_flat, wc, G, (long)N * N, v, stream);
    if (rc) return rc;
    for (int sp = 0; sp < 2; ++sp) {
      double* ws_ = sp == 0 ? walkers_a : walkers_b;
      const int ns = sp == 0 ? na : nb;
      if (!ws_) continue;
      // in place per chunk: the chunk is read completely into wtmp before it is overwritten
      double* wc_ptr = ws_ + 2 * (size_t)w0 * N * ns;
      if ((rc = trot_one_body_apply(expm, wc_ptr, wc, N, ns, wtmp, stream))) return rc;
      if ((rc = trot_taylor_apply(v, wtmp, wc, N, ns, sqrt_dt, n_terms, tws, tws_bytes, stream))) return rc;
      if ((rc = trot_one_body_apply(expm, wtmp, wc, N, ns, wc_ptr, stream))) return rc;
    }
  }
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// phaseless weight update.  scalars[0] = pop_control_ene_shift (in/out),
// scalars[1] = e_estimate (in); nodes[0] accumulates node encounters.
// ---------------------------------------------------------------------------
int trot_phaseless_update(double* weights, const double* ov_old, const double* ov_new, const double* shift_term,
                          const double* fb_term, double* scalars, long long* nodes, int W, double dt, double h0p_re,
                          double h0p_im, double w_floor, double w_cap, double damping, void* stream) {
  if (!weights || !ov_old || !ov_new || !shift_term || !fb_term || !scalars || !nodes)
    return fail(TROT_ERR_INVALID, "phaseless_update: null pointer");
  if (W < 1 || !(dt > 0)) return fail(TROT_ERR_INVALID, "phaseless_update: bad sizes");
  HIPCHK(launch_weight_update(weights, ov_old, ov_new, shift_term, fb_term, scalars, nodes, W, dt, sqrt(dt), h0p_re,
                              h0p_im, w_floor, w_cap, damping, (hipStream_t)stream));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// local energy from cached half Green's functions
// ---------------------------------------------------------------------------
static const int kEnergyGB = 32;

size_t trot_local_energy_workspace(int W, int G) {
  const int nblk = (G + kEnergyGB - 1) / kEnergyGB;
  return align256((size_t)W * (nblk > 0 ? nblk : 1) * 16);
}

int trot_local_energy(const double* rot_chol_a, const double* rot_chol_b, const double* green_t_a,
                      const double* green_t_b, const double* rot_h1t_a, const double* rot_h1t_b, int W, int N,
                      int na, int nb, int G, double h0, double* energy, void* ws, size_t ws_bytes, void* stream) {
  if (!green_t_a || !rot_h1t_a || !energy || (G > 0 && !rot_chol_a))
    return fail(TROT_ERR_INVALID, "local_energy: null pointer");
  if (green_t_b && (!rot_h1t_b || (G > 0 && !rot_chol_b)))
    return fail(TROT_ERR_INVALID, "local_energy: second spin incomplete");
  if (W < 1 || N < 1 || na < 1 || G < 0) return fail(TROT_ERR_INVALID, "local_energy: bad sizes");
  if (na > TROT_MAX_NOCC_BIG || (green_t_b && nb > TROT_MAX_NOCC_BIG))
    return fail(TROT_ERR_UNSUPPORTED, "local_energy: nocc > 256");
  if (ws_bytes < trot_local_energy_workspace(W, G)) return fail(TROT_ERR_WORKSPACE, "local_energy: workspace");
  HIPCHK(launch_energy(rot_chol_a, rot_chol_b, green_t_a, green_t_b, rot_h1t_a, rot_h1t_b, na, nb, N, G, W, h0,
                       kEnergyGB, (double*)ws, energy, (hipStream_t)stream));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// block energy tail: clamp, weighted mean, EMA of e_estimate (scalars[1])
// ---------------------------------------------------------------------------
int trot_block_energy(const double* energy, double* weights, double* scalars, double* out, int W, double dt,
                      double shift_ema, void* stream) {
  if (!energy || !weights || !scalars || !out) return fail(TROT_ERR_INVALID, "block_energy: null pointer");
  if (W < 1 || !(dt > 0)) return fail(TROT_ERR_INVALID, "block_energy: bad sizes");
  HIPCHK(launch_block_energy(energy, weights, scalars, out, W, sqrt(2.0 / dt), shift_ema, (hipStream_t)stream));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// orthonormalisation by two passes of Cholesky-QR: Q = w R^-1 with R the
// Cholesky factor of w^H w (real positive diagonal, the phase convention of
// trot/walkers.py:137-148).  A walker whose Cholesky breaks down in either pass
// (ill-conditioned: a pivot not finite or below 1e-8 of its Gram diagonal) is
// redone from its original columns by Gram-Schmidt with reorthogonalisation
// before the second pass's product, which then multiplies it by the identity.
// ---------------------------------------------------------------------------
size_t trot_orthonormalize_workspace(int W, int N, int n) {
  return 2 * align256((size_t)W * n * n * 16) + align256((size_t)W * N * n * 16) + align256((size_t)W * 4);
}

int trot_orthonormalize(double* walkers, int W, int N, int n, void* ws, size_t ws_bytes, void* stream) {
  return trot_orthogonalize(walkers, W, N, n, nullptr, ws, ws_bytes, stream);
}

int trot_orthogonalize(double* walkers, int W, int N, int n, double* det_r, void* ws, size_t ws_bytes, void* stream) {
  if (!walkers) return fail(TROT_ERR_INVALID, "orthonormalize: null pointer");
  if (W < 1 || N < 1 || n < 1 || n > N) return fail(TROT_ERR_INVALID, "orthonormalize: bad sizes");
  if (n > TROT_MAX_NOCC_BIG) return fail(TROT_ERR_UNSUPPORTED, "orthonormalize: nocc > 256");
  if (n > TROT_MAX_NOCC) {  // Gram-Schmidt with reorthogonalisation for every walker, in place
    HIPCHK(launch_qr_fallback(walkers, walkers, nullptr, N, n, W, nullptr, det_r, (hipStream_t)stream));
    return TROT_OK;
  }
  Carve c{(char*)ws, ws_bytes, 0};
  double* gram = c.take<double>((size_t)W * n * n * 2);
  double* rinv = c.take<double>((size_t)W * n * n * 2);
  double* tmp = c.take<double>((size_t)W * N * n * 2);
  int* flags = c.take<int>((size_t)W);
  if (!gram || !rinv || !tmp || !flags) return fail(TROT_ERR_WORKSPACE, "orthonormalize: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  double* src = walkers;
  double* dst = tmp;
  for (int pass = 0; pass < 2; ++pass) {
    GemmP p{};
    p.A = src; p.lda = n; p.sA = (long)N * n;
    p.B = src; p.ldb = n; p.sB = (long)N * n;
    p.C = gram; p.ldc = n; p.sC = (long)n * n;
    p.M = n; p.N = n; p.K = N; p.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, true, true, false, true, true, EPI_STORE>(p, W, s)));
    HIPCHK(launch_chol_rinv(gram, rinv, n, W, s, det_r, pass, flags));
    // the walkers are still intact here: broken ones get Q in tmp and R^-1 = I for the second product
    if (pass == 1) HIPCHK(launch_qr_fallback(walkers, tmp, flags, N, n, W, rinv, det_r, s));
    GemmP q{};
    q.A = src; q.lda = n; q.sA = (long)N * n;
    q.B = rinv; q.ldb = n; q.sB = (long)n * n;
    q.C = dst; q.ldc = n; q.sC = (long)N * n;
    q.M = N; q.N = n; q.K = n; q.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, true, true, true, true, false, EPI_STORE>(q, W, s)));
    src = dst;
    dst = walkers;
  }
  return TROT_OK;
}

size_t trot_qr_norm_workspace(int W, int N, int n) {
  return align256((size_t)W * n * n * 16) + align256((size_t)W * N * n * 16) + align256((size_t)W * 4);
}

int trot_qr_norm(const double* walkers, int W, int N, int n, double* det_r, void* ws, size_t ws_bytes, void* stream) {
  if (!walkers || !det_r) return fail(TROT_ERR_INVALID, "qr_norm: null pointer");
  if (W < 1 || N < 1 || n < 1 || n > N) return fail(TROT_ERR_INVALID, "qr_norm: bad sizes");
  if (n > TROT_MAX_NOCC_BIG) return fail(TROT_ERR_UNSUPPORTED, "qr_norm: nocc > 256");
  Carve c{(char*)ws, ws_bytes, 0};
  double* gram = c.take<double>((size_t)W * n * n * 2);
  double* q = c.take<double>((size_t)W * N * n * 2);
  int* flags = c.take<int>((size_t)W);
  if (!gram || !q || !flags) return fail(TROT_ERR_WORKSPACE, "qr_norm: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if (n > TROT_MAX_NOCC) {  // |det R| from Gram-Schmidt for every walker
    HIPCHK(launch_qr_fallback(walkers, q, nullptr, N, n, W, nullptr, det_r, s));
    return TROT_OK;
  }
  GemmP p{};
  p.A = walkers; p.lda = n; p.sA = (long)N * n;
  p.B = walkers; p.ldb = n; p.sB = (long)N * n;
  p.C = gram; p.ldc = n; p.sC = (long)n * n;
  p.M = n; p.N = n; p.K = N; p.ar = 1.0;
  HIPCHK((launch_gemm_f64<64, 64, true, true, false, true, true, EPI_STORE>(p, W, s)));
  HIPCHK(launch_chol_rinv(gram, nullptr, n, W, s, det_r, 0, flags));
  HIPCHK(launch_qr_fallback(walkers, q, flags, N, n, W, nullptr, det_r, s));  // |det R| of broken walkers
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// stochastic reconfiguration (comb) and walker gather
// ---------------------------------------------------------------------------
size_t trot_sr_workspace(int W) { return align256((size_t)W * 8); }

int trot_sr_comb(const double* weights, double zeta, int W, int* idx, double* weights_new, void* ws,
                 size_t ws_bytes, void* stream) {
  if (!weights || !idx || !weights_new) return fail(TROT_ERR_INVALID, "sr_comb: null pointer");
  if (W < 1) return fail(TROT_ERR_INVALID, "sr_comb: bad sizes");
  if (ws_bytes < trot_sr_workspace(W)) return fail(TROT_ERR_WORKSPACE, "sr_comb: workspace too small");
  HIPCHK(launch_sr_index(weights, (double*)ws, idx, weights_new, W, zeta, nullptr, (hipStream_t)stream));
  return TROT_OK;
}

int trot_sr_comb_dev(const double* weights, const double* zeta, int W, int* idx, double* weights_new, void* ws,
                     size_t ws_bytes, void* stream) {
  if (!weights || !zeta || !idx || !weights_new) return fail(TROT_ERR_INVALID, "sr_comb_dev: null pointer");
  if (W < 1) return fail(TROT_ERR_INVALID, "sr_comb_dev: bad sizes");
  if (ws_bytes < trot_sr_workspace(W)) return fail(TROT_ERR_WORKSPACE, "sr_comb_dev: workspace too small");
  HIPCHK(launch_sr_index(weights, (double*)ws, idx, weights_new, W, 0.0, zeta, (hipStream_t)stream));
  return TROT_OK;
}

int trot_sr_comb_dev_noalloc(const double* weights, const double* zeta, int W, int* idx, double* weights_new,
                             void* stream) {
  if (!weights || !zeta || !idx || !weights_new) return fail(TROT_ERR_INVALID, "sr_comb_dev_noalloc: null pointer");
  if (W < 1) return fail(TROT_ERR_INVALID, "sr_comb_dev_noalloc: bad sizes");
  if (W > TROT_SR_NOALLOC_MAX_W)
    return fail(TROT_ERR_UNSUPPORTED, "sr_comb_dev_noalloc: W above TROT_SR_NOALLOC_MAX_W (the static scratch)");
  double* cw = nullptr;  // the kernel writes all W cumulative weights before reading any: nothing to clear
  const int rc = scratch_on_current_device(&cw);
  if (rc != TROT_OK) return rc;
  HIPCHK(launch_sr_index(weights, cw, idx, weights_new, W, 0.0, zeta, (hipStream_t)stream));
  return TROT_OK;
}

int trot_gather_rows(const double* src, double* dst, const int* idx, int W, long row_doubles, void* stream) {
  if (!src || !dst || !idx || src == dst) return fail(TROT_ERR_INVALID, "gather_rows: bad pointers");
  if (W < 0 || row_doubles < 0) return fail(TROT_ERR_INVALID, "gather_rows: bad sizes");
  HIPCHK(launch_gather_rows(src, dst, idx, W, row_doubles, (hipStream_t)stream));
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// context builds (trot/prop/chol_afqmc_ops.py:74-127, trot/meas/rhf.py:308-319)
// ---------------------------------------------------------------------------
int trot_build_rot_chol(const double* chol, const double* trial, int G, int N, int n, double* rot_chol,
                        double* rot_chol_t, void* stream) {
  if (!chol || !trial) return fail(TROT_ERR_INVALID, "build_rot_chol: null pointer");
  if (G < 1 || N < 1 || n < 1 || G > 65535) return fail(TROT_ERR_INVALID, "build_rot_chol: bad sizes");
  hipStream_t s = (hipStream_t)stream;
  if (rot_chol) {  // rot_chol[g] = C^T L_g  (n x N)
    GemmP p{};
    p.A = trial; p.lda = n; p.sA = 0;
    p.B = chol; p.ldb = N; p.sB = (long)N * N;
    p.C = rot_chol; p.ldc = N; p.sC = (long)n * N;
    p.M = n; p.N = N; p.K = N; p.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, false, false, false, true, false, EPI_STORE>(p, G, s)));
  }
  if (rot_chol_t) {  // rot_chol_t[g] = L_g^T C  (N x n)
    GemmP p{};
    p.A = chol; p.lda = N; p.sA = (long)N * N;
    p.B = trial; p.ldb = n; p.sB = 0;
    p.C = rot_chol_t; p.ldc = n; p.sC = (long)N * n;
    p.M = N; p.N = n; p.K = N; p.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, false, false, false, true, false, EPI_STORE>(p, G, s)));
  }
  return TROT_OK;
}

size_t trot_build_prop_workspace(int G, int N) {
  const int gc = G < 64 ? G : 64;
  return align256((size_t)gc * N * N * 8);
}

// mf (G complex) = i tr(L_g dm); mfl (N x N complex) = sum_g mf_g L_g;
// v0m (N x N real) = 1/2 sum_g L_g L_g.
int trot_build_prop_tensors(const double* chol, const double* dm, int G, int N, double* mf, double* mfl,
                            double* v0m, void* ws, size_t ws_bytes, void* stream) {
  if (!chol || !dm || !mf || !mfl || !v0m) return fail(TROT_ERR_INVALID, "build_prop_tensors: null pointer");
  if (G < 1 || N < 1) return fail(TROT_ERR_INVALID, "build_prop_tensors: bad sizes");
  if (ws_bytes < trot_build_prop_workspace(G, N))
    return fail(TROT_ERR_WORKSPACE, "build_prop_tensors: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(launch_mf_trace(chol, dm, G, N, mf, s));
  {
    GemmP p{};
    p.A = mf; p.lda = G;
    p.B = chol; p.ldb = (long)N * N;
    p.C = mfl; p.ldc = (long)N * N;
    p.M = 1; p.N = N * N; p.K = G; p.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, true, false, true, true, false, EPI_STORE>(p, 1, s)));
  }
  const int gc = G < 64 ? G : 64;
  double* tmp = (double*)ws;
  for (int g0 = 0; g0 < G; g0 += gc) {
    const int nb = (G - g0) < gc ? (G - g0) : gc;
    GemmP p{};
    p.A = chol + (size_t)g0 * N * N; p.lda = N; p.sA = (long)N * N;
    p.B = chol + (size_t)g0 * N * N; p.ldb = N; p.sB = (long)N * N;
    p.C = tmp; p.ldc = N; p.sC = (long)N * N;
    p.M = N; p.N = N; p.K = N; p.ar = 1.0;
    HIPCHK((launch_gemm_f64<64, 64, false, false, true, true, false, EPI_STORE>(p, nb, s)));
    HIPCHK(launch_sum_batches(tmp, v0m, nb, (long)N * N, 0.5, g0 > 0, s));
  }
  return TROT_OK;
}

// ---------------------------------------------------------------------------
// grouped launches: one launch per operation over many walker sets (the batching form)
// ---------------------------------------------------------------------------
int trot_force_bias_slices(int W, int G, long Nn) {
  if (W < 1 || G < 1 || Nn < 1) return fail(TROT_ERR_INVALID, "force_bias_slices: bad sizes");
  return force_bias_slices(W, G, Nn);
}

// The GEMM of trot_force_bias_splitk as one problem of a grouped launch (op TROT_GOP_FORCE_BIAS): with one
// slice (trot_force_bias_slices) it writes fb = alpha acc (+ fb if accumulate) into out; with S > 1 it
// writes the S FP64 partial products into out (S W G complex) for a trot_sum_problem to add in slice
// order, as trot_force_bias_splitk does.  Returns the workgroup count.
long trot_force_bias_problem(const double* green_t, const double* rot_chol_t, int W, int G, long Nn, double alpha,
                             int accumulate, double* out, void* desc) {
  if (!green_t || !rot_chol_t || !out || !desc) return fail(TROT_ERR_INVALID, "force_bias_problem: null pointer");
  if (W < 1 || G < 1 || Nn < 1 || Nn >= (1L << 31)) return fail(TROT_ERR_INVALID, "force_bias_problem: bad sizes");
  const int S = force_bias_slices(W, G, Nn);
  GemmP p{};
  p.A = green_t; p.lda = Nn;
  p.B = rot_chol_t; p.ldb = Nn;
  p.C = out; p.ldc = G;
  p.M = W; p.N = G;
  if (S == 1) {
    p.K = (int)Nn; p.ar = alpha; p.beta = accumulate ? 1.0 : 0.0;
  } else {
    const long ks = (Nn + S - 1) / S;
    p.sA = ks; p.sB = ks; p.sC = (long)W * G;
    p.K = (int)ks; p.Klast = (int)(Nn - (S - 1) * ks); p.ar = 1.0;
  }
  *reinterpret_cast<GemmP*>(desc) = p;
  return gemm_f64_workgroups<64, 64>(p, S);
}

size_t trot_sum_problem_bytes(void) { return sizeof(SumProb); }
size_t trot_shift_problem_bytes(void) { return sizeof(ShiftProb); }
size_t trot_update_problem_bytes(void) { return sizeof(UpdProb); }

// dst (len doubles) = (accumulate ? dst : 0) + scale sum_{b < nb} src[b len ...]: the slice sum of
// trot_force_bias_splitk.  Returns the workgroup count.
long trot_sum_problem(const double* src, double* dst, int nb, long len, double scale, int accumulate, void* desc) {
  if (!src || !dst || !desc || nb < 1 || len < 1) return fail(TROT_ERR_INVALID, "sum_problem: bad arguments");
  *reinterpret_cast<SumProb*>(desc) = SumProb{src, dst, len, nb, accumulate, scale};
  return (len + 255) / 256;
}

int trot_sum_grouped(const void* probs, const int* start, int nprob, int total_blocks, void* stream) {
  if (!probs || !start || nprob < 1 || total_blocks < 1) return fail(TROT_ERR_INVALID, "sum_grouped: bad arguments");
  HIPCHK(launch_sum_grouped((const SumProb*)probs, start, nprob, total_blocks, (hipStream_t)stream));
  return TROT_OK;
}

// trot_field_shift of one walker set; returns W (workgroups).
long trot_shift_problem(const double* fields, const double* fb, const double* mf, double sqrt_dt, int W, int G,
                        double* xbar, double* shift_term, double* fb_term, void* desc) {
  if (!fields || !fb || !mf || !xbar || !shift_term || !fb_term || !desc)
    return fail(TROT_ERR_INVALID, "shift_problem: null pointer");
  if (W < 1 || G < 1) return fail(TROT_ERR_INVALID, "shift_problem: bad sizes");
  *reinterpret_cast<ShiftProb*>(desc) = ShiftProb{fields, fb, mf, xbar, shift_term, fb_term, sqrt_dt, G};
  return W;
}

int trot_shift_grouped(const void* probs, const int* start, int nprob, int total_w, void* stream) {
  if (!probs || !start || nprob < 1 || total_w < 1) return fail(TROT_ERR_INVALID, "shift_grouped: bad arguments");
  HIPCHK(launch_shift_grouped((const ShiftProb*)probs, start, nprob, total_w, (hipStream_t)stream));
  return TROT_OK;
}

// trot_overlap_combine + trot_phaseless_update + the copy of the new overlaps over the old ones, for one
// walker set (det_b NULL for restricted walkers).
int trot_update_problem(const double* det_a, const double* det_b, double* overlaps, const double* shift_term,
                        const double* fb_term, double* weights, double* scalars, long long* nodes, int W, double dt,
                        double h0p_re, double h0p_im, double w_floor, double w_cap, double damping, void* desc) {
  if (!det_a || !overlaps || !shift_term || !fb_term || !weights || !scalars || !nodes || !desc)
    return fail(TROT_ERR_INVALID, "update_problem: null pointer");
  if (W < 1 || !(dt > 0)) return fail(TROT_ERR_INVALID, "update_problem: bad sizes");
  *reinterpret_cast<UpdProb*>(desc) = UpdProb{det_a, det_b, overlaps, shift_term, fb_term, weights, scalars, nodes,
                                               W, dt, sqrt(dt), h0p_re, h0p_im, w_floor, w_cap, damping};
  return TROT_OK;
}

int trot_update_grouped(const void* probs, int nprob, void* stream) {
  if (!probs || nprob < 1) return fail(TROT_ERR_INVALID, "update_grouped: bad arguments");
  if (nprob > 65535 * 256) return fail(TROT_ERR_UNSUPPORTED, "update_grouped: too many sets");
  HIPCHK(launch_update_grouped((const UpdProb*)probs, nprob, (hipStream_t)stream));
  return TROT_OK;
}

size_t trot_gemm_problem_bytes(void) { return sizeof(GemmP); }
size_t trot_inv_problem_bytes(void) { return sizeof(InvProb); }
size_t trot_copy_problem_bytes(void) { return sizeof(CopyProb); }

// Fills desc (trot_gemm_problem_bytes()) with the problem of one operation over W walkers and returns the
// number of workgroups it takes in trot_gemm_f64_grouped, or a negative status.  Arguments, by op:
//   TROT_GOP_ONE_BODY  a = exp(-dt/2 h1_eff) (N x N), b = w_in, c = w_out
//   TROT_GOP_TAYLOR    a = V (sym-block), b = previous term, c = next term, d = running sum, coef = sqrt(dt)/k
//                      (32 x 64 tiles; TROT_GOP_TAYLOR_NARROW 32 x 32, TROT_GOP_TAYLOR_64x64 and _64x32 the
//                      other shapes of trot_autotune_taylor; every shape gives the same bits)
//   TROT_GOP_ONE_BODY_NARROW  TROT_GOP_ONE_BODY on 64 x 32 tiles (n <= 32), the same bits
//   TROT_GOP_OVERLAP   a = trial (N x n), b = walkers, c = C^T w (W, n, n)
//   TROT_GOP_GREEN     a = walkers, b = (C^T w)^-1, c = green_t (W, N, n)
//   TROT_GOP_VHS       a = xbar (W, G), b = chol (G, N2), c = vhs (W, N2); n = G, N = N2 (N2 < 2^31)
long trot_gemm_problem(int op, const double* a, const double* b, double* c, double* d, int W, int N, int n,
                       double coef, void* desc) {
  if (!a || !b || !c || !desc || W < 1 || N < 1 || n < 1) return fail(TROT_ERR_INVALID, "gemm_problem: bad arguments");
  GemmP p{};
  long wg = 0;
  switch (op) {
    case TROT_GOP_ONE_BODY: p = problem_one_body(a, b, c, N, n); wg = gemm_f64_workgroups<64, 64>(p, W); break;
    case TROT_GOP_TAYLOR:
      if (!d) return fail(TROT_ERR_INVALID, "gemm_problem: Taylor needs the running sum");
      p = problem_taylor_symblk(a, b, c, d, N, n, coef);
      wg = gemm_f64_workgroups<32, 64>(p, W);
      break;
    case TROT_GOP_TAYLOR_NARROW:
      if (!d) return fail(TROT_ERR_INVALID, "gemm_problem: Taylor needs the running sum");
      p = problem_taylor_symblk(a, b, c, d, N, n, coef);
      wg = gemm_f64_workgroups<32, 32>(p, W);
      break;
    case TROT_GOP_TAYLOR_64x64:
    case TROT_GOP_TAYLOR_64x32:
      if (!d) return fail(TROT_ERR_INVALID, "gemm_problem: Taylor needs the running sum");
      p = problem_taylor_symblk(a, b, c, d, N, n, coef);
      wg = op == TROT_GOP_TAYLOR_64x64 ? gemm_f64_workgroups<64, 64>(p, W) : gemm_f64_workgroups<64, 32>(p, W);
      break;
    case TROT_GOP_ONE_BODY_NARROW: p = problem_one_body(a, b, c, N, n); wg = gemm_f64_workgroups<64, 32>(p, W); break;
    case TROT_GOP_OVERLAP: p = problem_overlap(a, b, c, N, n); wg = gemm_f64_workgroups<64, 64>(p, W); break;
    case TROT_GOP_GREEN: p = problem_green(a, b, c, N, n); wg = gemm_f64_workgroups<64, 64>(p, W); break;
    case TROT_GOP_VHS: p = problem_vhs(a, b, c, n, N, W); wg = gemm_f64_workgroups<64, 64>(p, 1); break;
    default: return fail(TROT_ERR_INVALID, "gemm_problem: unknown operation");
  }
  *reinterpret_cast<GemmP*>(desc) = p;
  return wg;
}

// probs: device array of nprob problems from trot_gemm_problem (all of operation op); start: device
// prefix sums of their workgroup counts (nprob + 1 ints, start[0] = 0, start[nprob] = total_wg).
int trot_gemm_f64_grouped(int op, const void* probs, const int* start, int nprob, int total_wg, void* stream) {
  if (!probs || !start || nprob < 1 || total_wg < 1) return fail(TROT_ERR_INVALID, "gemm_f64_grouped: bad arguments");
  const GemmP* P = (const GemmP*)probs;
  hipStream_t s = (hipStream_t)stream;
  switch (op) {
    case TROT_GOP_ONE_BODY:
      HIPCHK((launch_gemm_f64_grouped<64, 64, false, true, true, true, false, EPI_STORE>(P, start, nprob, total_wg, s)));
      break;
    case TROT_GOP_TAYLOR:
      HIPCHK((launch_gemm_f64_grouped<32, 64, true, true, true, true, false, EPI_TAYLOR, true>(P, start, nprob, total_wg,
                                                                                           s)));
      break;
    case TROT_GOP_TAYLOR_NARROW:
      HIPCHK((launch_gemm_f64_grouped<32, 32, true, true, true, true, false, EPI_TAYLOR, true>(P, start, nprob, total_wg,
                                                                                           s)));
      break;
    case TROT_GOP_TAYLOR_64x64:
      HIPCHK((launch_gemm_f64_grouped<64, 64, true, true, true, true, false, EPI_TAYLOR, true>(P, start, nprob, total_wg,
                                                                                           s)));
      break;
    case TROT_GOP_TAYLOR_64x32:
      HIPCHK((launch_gemm_f64_grouped<64, 32, true, true, true, true, false, EPI_TAYLOR, true>(P, start, nprob, total_wg,
                                                                                           s)));
      break;
    case TROT_GOP_ONE_BODY_NARROW:
      HIPCHK((launch_gemm_f64_grouped<64, 32, false, true, true, true, false, EPI_STORE>(P, start, nprob, total_wg, s)));
      break;
    case TROT_GOP_OVERLAP:
      HIPCHK((launch_gemm_f64_grouped<64, 64, false, true, false, true, false, EPI_STORE>(P, start, nprob, total_wg, s)));
      break;
    case TROT_GOP_GREEN:
      HIPCHK((launch_gemm_f64_grouped<64, 64, true, true, true, true, false, EPI_STORE>(P, start, nprob, total_wg, s)));
      break;
    case TROT_GOP_VHS:
      HIPCHK((launch_gemm_f64_grouped<64, 64, true, false, true, true, false, EPI_STORE>(P, start, nprob, total_wg, s)));
      break;
    case TROT_GOP_FORCE_BIAS:  // problems from trot_force_bias_problem
      HIPCHK((launch_gemm_f64_grouped<64, 64, true, false, true, false, false, EPI_STORE>(P, start, nprob, total_wg,
                                                                                        s)));
      break;
    default: return fail(TROT_ERR_INVALID, "gemm_f64_grouped: unknown operation");
  }
  return TROT_OK;
}

long trot_inv_problem(const double* m, double* minv, double* det, int W, int n, void* desc) {
  if (!m || !minv || !det || !desc || W < 1 || n < 1) return fail(TROT_ERR_INVALID, "inv_problem: bad arguments");
  if (n > TROT_MAX_NOCC) return fail(TROT_ERR_UNSUPPORTED, "inv_problem: nocc > 64");
  *reinterpret_cast<InvProb*>(desc) = InvProb{m, minv, det, n};
  return W;
}

int trot_inv_det_grouped(const void* probs, const int* start, int nprob, int total_w, void* stream) {
  if (!probs || !start || nprob < 1 || total_w < 1) return fail(TROT_ERR_INVALID, "inv_det_grouped: bad arguments");
  HIPCHK(launch_inv_det_grouped((const InvProb*)probs, start, nprob, total_w, (hipStream_t)stream));
  return TROT_OK;
}

int trot_copy_problem(const void* src, void* dst, long bytes, void* desc) {
  if (!src || !dst || !desc || bytes < 0 || bytes % 16) return fail(TROT_ERR_INVALID, "copy_problem: bad arguments");
  *reinterpret_cast<CopyProb*>(desc) = CopyProb{src, dst, bytes};
  return TROT_OK;
}

int trot_copy_grouped(const void* probs, int nprob, long max_bytes, void* stream) {
  if (!probs || nprob < 1 || max_bytes < 0) return fail(TROT_ERR_INVALID, "copy_grouped: bad arguments");
  HIPCHK(launch_copy_grouped((const CopyProb*)probs, nprob, max_bytes, (hipStream_t)stream));
  return TROT_OK;
}

}  // extern "C"
