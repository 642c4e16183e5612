// Per-walker FP64 kernels of the phaseless AFQMC step (gfx950).
//
//   inv_det_kernel        m -> (m^-1, det m): overlap and half Green's function
//                         (trot/trial/rhf.py:45-47, trot/meas/rhf.py:21-25)
//   shift_kernel          force bias -> shifted fields, shift and bias terms
//                         (trot/prop/afqmc.py:99-103)
//   weight_update_kernel  phaseless importance, masking, population shift
//                         (trot/prop/afqmc.py:108-134)
//   chol_rinv_kernel      Gram matrix -> R^-1 of its Cholesky factor (CholQR), breakdown flags
//                         (trot/walkers.py:137-179)
//   qr_fallback_kernel    Gram-Schmidt QR (twice) of the walkers whose CholQR broke down
//   sr_index_kernel       comb stochastic reconfiguration indices
//                         (trot/walkers.py:249-293)
//   gather_rows_kernel    walker / cache gather after SR
#include "common.hpp"
#include "kernels.hpp"

// ---------------------------------------------------------------------------
// Gauss-Jordan inversion with partial pivoting (|re|+|im| pivot, first index on
// ties, as LAPACK's izamax), det = prod(pivots) * (-1)^swaps.  One workgroup of
// four waves per walker, n <= 64; the matrix is padded to 64 x 64 with an
// identity block and held in registers: element (i, j) in lane i of wave j / 16,
// register j % 16.  Rows are never moved: each lane keeps the position its row
// would have after LAPACK's swaps (lp), the pivot search runs over lp >= k with
// ties to the smallest lp; the pivot column reaches every lane through LDS
// (double-buffered by step parity, so one barrier per step suffices) and each
// wave takes its columns of the pivot row from lane pr by shuffles.  The
// arithmetic per row is the row-swapping algorithm's; only the storage differs.
// ---------------------------------------------------------------------------
struct InvShared {
  cplx colk[2][64];  // column k of every (physical) row, before the step
  int s_piv[2];
  int perm[64];      // position swapped with k at step k (LAPACK ipiv)
  int outcol[64];
};

// One walker's n x n matrix at mw -> inverse at ow, determinant at detw (one complex).
__device__ __forceinline__ void inv_det_one(const double* __restrict__ mw, double* __restrict__ ow,
                                            double* __restrict__ detw, int n, InvShared& sh) {
  cplx(&colk)[2][64] = sh.colk;
  int(&s_piv)[2] = sh.s_piv;
  int(&perm)[64] = sh.perm;
  int(&outcol)[64] = sh.outcol;
  const int t = threadIdx.x, lane = t & 63, wq = t >> 6;
  cplx a[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int j = wq * 16 + r;
    a[r] = (lane < n && j < n) ? cload(mw, (long)lane * n + j) : cplx{lane == j ? 1.0 : 0.0, 0.0};
  }
  int lp = lane;  // position of this row after the swaps so far (identical in all four waves)
  cplx dacc = {1.0, 0.0};
  for (int k = 0; k < n; ++k) {
    const int kw = k >> 4, kr = k & 15, pb = k & 1;
    if (wq == kw) {
      cplx x = a[0];
#pragma unroll
      for (int r = 1; r < 16; ++r)
        if (r == kr) x = a[r];
      colk[pb][lane] = x;
      const double c = cabs1(x);
      double v = (lp >= k && lane < n && c == c) ? c : -1.0;  // NaN never wins
      int key = lp;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off);
        const int k2 = __shfl_xor(key, off);
        if (v2 > v || (v2 == v && k2 < key)) {
          v = v2;
          key = k2;
        }
      }
      // the winning row is the lane whose position is key (positions are distinct); a column of NaNs
      // keeps the row at position k
      const unsigned long long at = __ballot(lp == (v < 0.0 ? k : key));
      const int bi = __ffsll((long long)at) - 1;
      if (lane == 0) s_piv[pb] = bi;
    }
    __syncthreads();
    const int pr = s_piv[pb];
    const int lpp = __shfl(lp, pr);  // position of the pivot row (>= k)
    if (t == 0) {
      perm[k] = lpp;
      dacc = cmul(dacc, colk[pb][pr]);
      if (lpp != k) dacc = {-dacc.re, -dacc.im};
    }
    if (lp == k) lp = lpp;  // the swap of positions k and lpp
    if (lane == pr) lp = k;
    const cplx piv = colk[pb][pr];
    const cplx ipiv = cinv(piv);
    // with u_j = a[pr][j] (1 at j = k): the pivot row becomes u ipiv, every other row i becomes
    // a[i][j] - a[i][k] ipiv u_j with column k taken as 0.  One update for all lanes:
    // a <- x - g u, (x, g) = (0, -ipiv) on the pivot row and (a with column k zeroed, a[i][k] ipiv) elsewhere.
    const bool is_p = lane == pr;
    const cplx fi = cmul(colk[pb][lane], ipiv);
    const cplx g = {is_p ? -ipiv.re : fi.re, is_p ? -ipiv.im : fi.im};
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int j = wq * 16 + r;
      cplx u = {__shfl(a[r].re, pr), __shfl(a[r].im, pr)};  // the pivot row, from lane pr of this wave
      if (j == k) u = {1.0, 0.0};
      const bool zero = is_p || j == k;
      const cplx x = {zero ? 0.0 : a[r].re, zero ? 0.0 : a[r].im};
      a[r] = csub(x, cmul(g, u));
    }
  }
  __syncthreads();
  if (t == 0) {  // undo the row swaps on the columns: column c of the inverse is column col[c] here
    for (int c = 0; c < 64; ++c) outcol[c] = c;
    for (int k = n - 1; k >= 0; --k) {
      const int pr = perm[k];
      const int x = outcol[k];
      outcol[k] = outcol[pr];
      outcol[pr] = x;
    }
  }
  __syncthreads();
  if (t < 64) colk[0][outcol[t]].re = (double)t;  // inverse map: here column j -> output column
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int j = wq * 16 + r;
    if (lp < n && j < n) cstore(ow, (long)lp * n + (int)colk[0][j].re, a[r]);
  }
  if (t == 0) cstore(detw, 0, dacc);
}

__global__ __launch_bounds__(256) void inv_det_kernel(const double* __restrict__ m, double* __restrict__ minv,
                                                      double* __restrict__ det, int n) {
  __shared__ InvShared sh;
  const long w = blockIdx.x;
  inv_det_one(m + 2 * w * n * n, minv + 2 * w * n * n, det + 2 * w, n, sh);
}

// Many walker sets in one launch: problem i owns workgroups start[i] .. start[i+1] - 1, one per walker.
__global__ __launch_bounds__(256) void inv_det_grouped_kernel(const InvProb* __restrict__ probs,
                                                              const int* __restrict__ start, int nprob) {
  __shared__ InvShared sh;
  const int g = blockIdx.x;
  int lo = 0, hi = nprob;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= g) lo = mid; else hi = mid;
  }
  const InvProb q = probs[lo];
  const long w = g - start[lo];
  inv_det_one(q.m + 2 * w * q.n * q.n, q.minv + 2 * w * q.n * q.n, q.det + 2 * w, q.n, sh);
}

// Many copies in one launch: grid (chunks of 64 KiB, problems); sizes in bytes, multiples of 16.
__global__ __launch_bounds__(256) void copy_grouped_kernel(const CopyProb* __restrict__ probs) {
  const CopyProb q = probs[blockIdx.y];
  const long n16 = q.bytes / 16;
  const long beg = (long)blockIdx.x * 4096;
  const long end = beg + 4096 < n16 ? beg + 4096 : n16;
  const double2* src = reinterpret_cast<const double2*>(q.src);
  double2* dst = reinterpret_cast<double2*>(q.dst);
  for (long e = beg + threadIdx.x; e < end; e += 256) dst[e] = src[e];
}

// ---------------------------------------------------------------------------
// overlap from determinants: RHF restricted walkers det^2, unrestricted det_a*det_b
// ---------------------------------------------------------------------------
__global__ void overlap_combine_kernel(const double* __restrict__ det_a, const double* __restrict__ det_b,
                                       double* __restrict__ ovlp, int W) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  const cplx a = cload(det_a, w);
  const cplx b = det_b ? cload(det_b, w) : a;
  cstore(ovlp, w, cmul(a, b));
}

// ---------------------------------------------------------------------------
// shifted fields.  fs = -sqrt(dt) (i fb - mf); xbar = x - fs;
// shift = sum_g xbar_g mf_g; fbt = sum_g (x_g fs_g - fs_g^2 / 2).
// One workgroup per walker, deterministic tree reduction.
// xbar32 (optional): complex64 copy for the FP32 / split-BF16 VHS.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void shift_one(const double* __restrict__ fields, const double* __restrict__ fb,
                                          const double* __restrict__ mf, double sqrt_dt, int G, long w,
                                          double* __restrict__ xbar, float* __restrict__ xbar32,
                                          double* __restrict__ shift, double* __restrict__ fbt, double* red) {
  const int t = threadIdx.x;
  double s_re = 0, s_im = 0, f_re = 0, f_im = 0;
  for (int g = t; g < G; g += 256) {
    const double x = fields[w * G + g];
    const cplx f = cload(fb, w * G + g);
    const cplx m = cload(mf, g);
    // i*fb - mf = (-f.im - m.re) + i (f.re - m.im)
    const cplx ifm = {-f.im - m.re, f.re - m.im};
    const cplx fsv = {-sqrt_dt * ifm.re, -sqrt_dt * ifm.im};
    const cplx xb = {x - fsv.re, -fsv.im};
    cstore(xbar, w * G + g, xb);
    if (xbar32) {
      *reinterpret_cast<float2*>(xbar32 + 2 * (w * G + g)) = make_float2((float)xb.re, (float)xb.im);
    }
    const cplx sm = cmul(xb, m);
    s_re += sm.re;
    s_im += sm.im;
    const cplx fs2 = cmul(fsv, fsv);
    f_re += x * fsv.re - 0.5 * fs2.re;
    f_im += x * fsv.im - 0.5 * fs2.im;
  }
  s_re = block_sum<256>(s_re, red);
  s_im = block_sum<256>(s_im, red);
  f_re = block_sum<256>(f_re, red);
  f_im = block_sum<256>(f_im, red);
  if (t == 0) {
    cstore(shift, w, {s_re, s_im});
    cstore(fbt, w, {f_re, f_im});
  }
}

__global__ __launch_bounds__(256) void shift_kernel(const double* __restrict__ fields, const double* __restrict__ fb,
                                                    const double* __restrict__ mf, double sqrt_dt, int G,
                                                    double* __restrict__ xbar, float* __restrict__ xbar32,
                                                    double* __restrict__ shift, double* __restrict__ fbt) {
  __shared__ double red[256];
  shift_one(fields, fb, mf, sqrt_dt, G, blockIdx.x, xbar, xbar32, shift, fbt, red);
}

// Problem index of block b from the prefix sums start[0..nprob] (start[0] = 0).
__device__ __forceinline__ int grouped_problem(const int* __restrict__ start, int nprob, int b) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One workgroup per (walker set, walker): shift_kernel's work for every set in one launch.
__global__ __launch_bounds__(256) void shift_grouped_kernel(const ShiftProb* __restrict__ probs,
                                                            const int* __restrict__ start, int nprob) {
  __shared__ double red[256];
  const int pi = grouped_problem(start, nprob, blockIdx.x);
  const ShiftProb& q = probs[pi];
  shift_one(q.fields, q.fb, q.mf, q.sqrt_dt, q.G, blockIdx.x - start[pi], q.xbar, nullptr, q.shift, q.fbt, red);
}

// sum_batches_kernel's work for every set: block b of set p covers elements (b - start[p]) * 256 ...
__global__ __launch_bounds__(256) void sum_grouped_kernel(const SumProb* __restrict__ probs,
                                                          const int* __restrict__ start, int nprob) {
  const int pi = grouped_problem(start, nprob, blockIdx.x);
  const SumProb& q = probs[pi];
  const long e = (long)(blockIdx.x - start[pi]) * 256 + threadIdx.x;
  if (e >= q.len) return;
  double acc = 0.0;
  for (int b = 0; b < q.nb; ++b) acc += q.src[(long)b * q.len + e];
  q.dst[e] = (q.accumulate ? q.dst[e] : 0.0) + q.scale * acc;
}

// ---------------------------------------------------------------------------
// phaseless weight update (single workgroup, deterministic).
// scal[0] = pop_control_ene_shift (in/out), scal[1] = e_estimate (in).
// nodes[0] += number of walkers with importance <= 0 after flooring.
// ---------------------------------------------------------------------------
// det_a (optional): the new overlap of walker w is computed here as det_a[w] det_b[w] (det_a[w]^2 without
// det_b), exactly as overlap_combine_kernel does, instead of read from ov_new; ov_out (optional) receives
// the new overlaps after the old ones were read (in place over ov_old is allowed: one thread per walker).
__device__ __forceinline__ void weight_update_wg(double* __restrict__ weights, const double* ov_old,
                                                 const double* __restrict__ ov_new, const double* __restrict__ det_a,
                                                 const double* __restrict__ det_b, double* ov_out,
                                                 const double* __restrict__ shift, const double* __restrict__ fbt,
                                                 double* __restrict__ scal, long long* __restrict__ nodes, int W,
                                                 double dt, double sqrt_dt, double h0p_re, double h0p_im,
                                                 double w_floor, double w_cap, double damping, double* red) {
  const int t = threadIdx.x;
  const double pop = scal[0];
  const double e_est = scal[1];
  double wsum = 0.0;
  double ncount = 0.0;
  for (int w = t; w < W; w += 1024) {
    cplx on;
    if (det_a) {
      const cplx a = cload(det_a, w);
      const cplx b = det_b ? cload(det_b, w) : a;
      on = cmul(a, b);
    } else {
      on = cload(ov_new, w);
    }
    const cplx oo = cload(ov_old, w);
    const cplx ratio = cmul(on, cinv(oo));
    const cplx st = cload(shift, w), ft = cload(fbt, w);
    const cplx ex = {-sqrt_dt * st.re + ft.re + dt * (pop + h0p_re), -sqrt_dt * st.im + ft.im + dt * h0p_im};
    const double mag = exp(ex.re);
    const cplx e1 = {mag * cos(ex.im), mag * sin(ex.im)};
    const cplx imp = cmul(e1, ratio);
    const double mag2 = exp(-sqrt_dt * st.re);
    const double ph2 = -sqrt_dt * st.im;
    const cplx e2 = {mag2 * cos(ph2), mag2 * sin(ph2)};
    const cplx z = cmul(e2, ratio);
    const double theta = atan2(z.im, z.re);
    double p = hypot(imp.re, imp.im) * cos(theta);
    if (!isfinite(p) || p < w_floor) p = 0.0;
    if (p <= 0.0) ncount += 1.0;
    if (p > w_cap) p = 0.0;
    double wn = weights[w] * p;
    if (wn > w_cap) wn = 0.0;
    weights[w] = wn;
    wsum += wn;
    if (ov_out) cstore(ov_out, w, on);
  }
  wsum = block_sum<1024>(wsum, red);
  ncount = block_sum<1024>(ncount, red);
  if (t == 0) {
    double avg = wsum / (double)W;
    if (avg < 1.0e-300) avg = 1.0e-300;
    scal[0] = e_est - damping * (log(avg) / dt);
    nodes[0] += (long long)ncount;
  }
}

__global__ __launch_bounds__(1024) void weight_update_kernel(
    double* __restrict__ weights, const double* __restrict__ ov_old, const double* __restrict__ ov_new,
    const double* __restrict__ shift, const double* __restrict__ fbt, double* __restrict__ scal,
    long long* __restrict__ nodes, int W, double dt, double sqrt_dt, double h0p_re, double h0p_im, double w_floor,
    double w_cap, double damping) {
  __shared__ double red[1024];
  weight_update_wg(weights, ov_old, ov_new, nullptr, nullptr, nullptr, shift, fbt, scal, nodes, W, dt, sqrt_dt,
                   h0p_re, h0p_im, w_floor, w_cap, damping, red);
}

// One workgroup per walker set: overlap combination, weight update and the new overlaps stored over the
// old ones (overlap_combine_kernel + weight_update_kernel + a copy, per set).
__global__ __launch_bounds__(1024) void update_grouped_kernel(const UpdProb* __restrict__ probs) {
  __shared__ double red[1024];
  const UpdProb& q = probs[blockIdx.x];
  weight_update_wg(q.weights, q.ovlp, nullptr, q.det_a, q.det_b, q.ovlp, q.shift, q.fbt, q.scal, q.nodes, q.W, q.dt,
                   q.sqrt_dt, q.h0p_re, q.h0p_im, q.w_floor, q.w_cap, q.damping, red);
}

// ---------------------------------------------------------------------------
// Cholesky of the Hermitian Gram matrix S = R^H R (R upper, real positive
// diagonal) and R^-1.  One workgroup per walker.  rinv gets R^-1 (n x n).
// ---------------------------------------------------------------------------
// dprod (optional): product of the Cholesky diagonal, prod_k L_kk = |det R| of w = Q R, written
// (accumulate = 0) or multiplied in (accumulate = 1, the second CholQR pass).  rinv null: diagonal only.
// flags (optional): 1 for a walker whose Cholesky breaks down, i.e. some pivot is not finite or falls below
// CHOLQR_PIVOT_TOL times its Gram diagonal entry (written by the first pass, or-ed in by the second); such a
// walker is redone by qr_fallback_kernel.
#define CHOLQR_PIVOT_TOL 1.0e-8
__global__ __launch_bounds__(256) void chol_rinv_kernel(const double* __restrict__ s, double* __restrict__ rinv,
                                                        int n, double* __restrict__ dprod, int accumulate,
                                                        int* __restrict__ flags) {
  extern __shared__ double sh[];
  cplx* a = reinterpret_cast<cplx*>(sh);
  cplx* x = a + n * n;
  const int t = threadIdx.x;
  const long w = blockIdx.x;
  const double* sw = s + 2 * w * n * n;
  for (int e = t; e < n * n; e += 256) a[e] = cload(sw, e);
  __syncthreads();
  double dp = 1.0;
  bool broken = false;
  for (int k = 0; k < n; ++k) {
    const double piv = a[k * n + k].re;
    // the Gram diagonal S_kk is still in global memory; pivot^2 >= sigma_min^2 while S_kk <= sigma_max^2
    broken = broken || !(piv > CHOLQR_PIVOT_TOL * sw[2 * ((long)k * n + k)]) || !isfinite(piv);
    const double d = sqrt(piv);
    dp *= d;
    __syncthreads();
    if (t == 0) a[k * n + k] = {d, 0.0};
    const double id = 1.0 / d;
    for (int j = k + 1 + t; j < n; j += 256) a[k * n + j] = cscale(a[k * n + j], id);
    __syncthreads();
    const int r = n - k - 1;
    for (int e = t; e < r * r; e += 256) {
      const int i = k + 1 + e / r, j = k + 1 + e % r;
      if (j < i) continue;
      a[i * n + j] = csub(a[i * n + j], cmul(cconj(a[k * n + i]), a[k * n + j]));
    }
    __syncthreads();
  }
  if (dprod && t == 0) dprod[w] = accumulate ? dprod[w] * dp : dp;
  if (flags && t == 0) flags[w] = (accumulate ? flags[w] : 0) | (broken ? 1 : 0);
  if (!rinv) return;
  // R X = I, X upper: column j by thread j (back substitution)
  for (int e = t; e < n * n; e += 256) x[e] = {0.0, 0.0};
  __syncthreads();
  for (int j = t; j < n; j += 256) {
    x[j * n + j] = {1.0 / a[j * n + j].re, 0.0};
    for (int i = j - 1; i >= 0; --i) {
      cplx acc = {0.0, 0.0};
      for (int l = i + 1; l <= j; ++l) acc = cadd(acc, cmul(a[i * n + l], x[l * n + j]));
      x[i * n + j] = cscale(acc, -1.0 / a[i * n + i].re);
    }
  }
  __syncthreads();
  double* ow = rinv + 2 * w * n * n;
  for (int e = t; e < n * n; e += 256) cstore(ow, e, x[e]);
}

// ---------------------------------------------------------------------------
// QR of the walkers whose Cholesky-QR broke down (flags[w] != 0; other workgroups return at once; null flags:
// every walker, the QR for n > 64, beyond Cholesky-QR's LDS-resident Gram matrix; q may equal src).  Classical
// Gram-Schmidt with reorthogonalisation ("twice is enough": orthonormal to rounding for any numerically
// full-rank w, unlike Cholesky-QR, which needs kappa(w) below ~1e8): column k is projected out of the
// orthonormal columns before it twice, then normalised, so R_kk = |residual| is real and >= 0, the phase
// convention of trot/walkers.py:141-148.  src (N x n, read) -> q (N x n, written); rinv (optional) set to the
// identity for the walker, so that a following q rinv product keeps Q; dprod (optional) = prod_k R_kk.
// A column in the span of the ones before it (exact rank deficiency) comes out as zeros and R_kk = 0.
// One workgroup of 4 waves per walker; coefficients one per wave with a shuffle reduction (fixed order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void qr_fallback_kernel(const double* __restrict__ src, double* __restrict__ q,
                                                          const int* __restrict__ flags, int N, int n,
                                                          double* __restrict__ rinv, double* __restrict__ dprod) {
  extern __shared__ double sh[];
  __shared__ double red[256];
  cplx* coef = reinterpret_cast<cplx*>(sh);
  const long w = blockIdx.x;
  if (flags && !flags[w]) return;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const long base = w * (long)N * n;
  for (long e = t; e < (long)N * n; e += 256) cstore(q, base + e, cload(src, base + e));
  __syncthreads();
  double dp = 1.0;
  for (int k = 0; k < n; ++k) {
    for (int pass = 0; pass < 2 && k > 0; ++pass) {
      for (int j = wave; j < k; j += 4) {
        cplx acc = {0.0, 0.0};
        for (int i = lane; i < N; i += 64) {
          const cplx qj = cload(q, base + (long)i * n + j), qk = cload(q, base + (long)i * n + k);
          acc = cadd(acc, cmul(cconj(qj), qk));
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
          acc.re += __shfl_xor(acc.re, off);
          acc.im += __shfl_xor(acc.im, off);
        }
        if (lane == 0) coef[j] = acc;
      }
      __syncthreads();
      for (int i = t; i < N; i += 256) {
        cplx x = cload(q, base + (long)i * n + k);
        for (int j = 0; j < k; ++j) x = csub(x, cmul(coef[j], cload(q, base + (long)i * n + j)));
        cstore(q, base + (long)i * n + k, x);
      }
      __syncthreads();
    }
    double part = 0.0;
    for (int i = t; i < N; i += 256) {
      const cplx x = cload(q, base + (long)i * n + k);
      part += x.re * x.re + x.im * x.im;
    }
    const double r = sqrt(block_sum<256>(part, red));
    dp *= r;
    const double inv = r > 0.0 ? 1.0 / r : 0.0;
    for (int i = t; i < N; i += 256) cstore(q, base + (long)i * n + k, cscale(cload(q, base + (long)i * n + k), inv));
    __syncthreads();
  }
  if (dprod && t == 0) dprod[w] = dp;
  if (rinv)
    for (int e = t; e < n * n; e += 256) cstore(rinv, w * (long)n * n + e, {e / n == e % n ? 1.0 : 0.0, 0.0});
}

// ---------------------------------------------------------------------------
// comb SR: cw = cumsum(|w|) (sequential, the order of numpy/XLA's scan),
// z_i = cw[-1] (i + zeta) / W, idx_i = searchsorted(cw, z_i, 'left') clamped to
// W-1 (gather semantics), new weights = cw[-1] / W.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void sr_index_kernel(const double* __restrict__ weights, double* __restrict__ cw,
                                                        int* __restrict__ idx, double* __restrict__ w_new, int W,
                                                        double zeta_v, const double* __restrict__ zeta_p) {
  const int t = threadIdx.x;
  const double zeta = zeta_p ? zeta_p[0] : zeta_v;
  if (t == 0) {
    double acc = 0.0;
    for (int i = 0; i < W; ++i) {
      acc += fabs(weights[i]);
      cw[i] = acc;
    }
  }
  __syncthreads();
  const double tot = cw[W - 1];
  for (int i = t; i < W; i += 1024) {
    const double z = tot * ((double)i + zeta) / (double)W;
    int lo = 0, hi = W;  // first index with cw[idx] >= z
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cw[mid] < z) lo = mid + 1; else hi = mid;
    }
    idx[i] = lo < W ? lo : W - 1;
  }
  __syncthreads();
  for (int i = t; i < W; i += 1024) w_new[i] = tot / (double)W;
}

// dst[i, :] = src[idx[i], :], rows of `row` doubles.  Grid (W).
__global__ __launch_bounds__(256) void gather_rows_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                                          const int* __restrict__ idx, long row) {
  const long i = blockIdx.x;
  const long j = idx[i];
  const double2* s = reinterpret_cast<const double2*>(src + j * row);
  double2* d = reinterpret_cast<double2*>(dst + i * row);
  const long n2 = row / 2;
  for (long e = threadIdx.x; e < n2; e += 256) d[e] = s[e];
  if ((row & 1) && threadIdx.x == 0) dst[i * row + row - 1] = src[j * row + row - 1];
}

// mf_g = i * sum_ij L_g[i,j] dm[j,i]  (trot/prop/chol_afqmc_ops.py:74-76), real L and dm.
__global__ __launch_bounds__(256) void mf_trace_kernel(const double* __restrict__ chol, const double* __restrict__ dm,
                                                       int N, double* __restrict__ mf) {
  __shared__ double red[256];
  const long g = blockIdx.x;
  const double* lg = chol + g * (long)N * N;
  double s = 0.0;
  for (long e = threadIdx.x; e < (long)N * N; e += 256) {
    const long i = e / N, j = e - (e / N) * N;
    s += lg[e] * dm[j * N + i];
  }
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) cstore(mf, g, {0.0, s});
}

// dst[e] (+)= scale * sum_b src[b * len + e], b ascending.
__global__ void sum_batches_kernel(const double* __restrict__ src, double* __restrict__ dst, int nb, long len,
                                   double scale, int accumulate) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += src[(long)b * len + e];
  dst[e] = (accumulate ? dst[e] : 0.0) + scale * s;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static hipError_t set_lds(const void* f, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

hipError_t launch_inv_det_grouped(const InvProb* probs, const int* start, int nprob, int total_w, hipStream_t s) {
  if (nprob <= 0 || total_w <= 0) return hipSuccess;
  hipLaunchKernelGGL(inv_det_grouped_kernel, dim3(total_w), dim3(256), 0, s, probs, start, nprob);
  return hipGetLastError();
}

hipError_t launch_copy_grouped(const CopyProb* probs, int nprob, long max_bytes, hipStream_t s) {
  if (nprob <= 0 || max_bytes <= 0) return hipSuccess;
  const long chunks = (max_bytes / 16 + 4095) / 4096;
  if (nprob > 65535 || chunks > 0x7fffffffL) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(copy_grouped_kernel, dim3((unsigned)chunks, (unsigned)nprob), dim3(256), 0, s, probs);
  return hipGetLastError();
}

#define INV_LDS_MAX 96  // inv_det_big_kernel keeps the matrix in LDS up to this n (147 KB)
__global__ void inv_det_big_kernel(const double* __restrict__ m, double* __restrict__ minv, double* __restrict__ det,
                                   int n);

hipError_t launch_inv_det(const double* m, double* minv, double* det, int n, int W, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  if (n < 1 || n > TROT_MAX_NOCC_BIG) return hipErrorInvalidValue;
  if (n <= TROT_MAX_NOCC) {
    hipLaunchKernelGGL(inv_det_kernel, dim3(W), dim3(256), 0, s, m, minv, det, n);
    return hipGetLastError();
  }
  const size_t lds = (n <= INV_LDS_MAX ? (size_t)n * n * 16 : 0) + (size_t)n * 16 + 256 * 8 + 256 * 4 + (size_t)n * 4;
  hipError_t e = set_lds((const void*)inv_det_big_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(inv_det_big_kernel, dim3(W), dim3(256), lds, s, m, minv, det, n);
  return hipGetLastError();
}

hipError_t launch_overlap_combine(const double* det_a, const double* det_b, double* ovlp, int W, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  hipLaunchKernelGGL(overlap_combine_kernel, dim3((W + 255) / 256), dim3(256), 0, s, det_a, det_b, ovlp, W);
  return hipGetLastError();
}

hipError_t launch_shift(const double* fields, const double* fb, const double* mf, double sqrt_dt, int W, int G,
                        double* xbar, float* xbar32, double* shift, double* fbt, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  hipLaunchKernelGGL(shift_kernel, dim3(W), dim3(256), 0, s, fields, fb, mf, sqrt_dt, G, xbar, xbar32, shift, fbt);
  return hipGetLastError();
}

hipError_t launch_weight_update(double* weights, const double* ov_old, const double* ov_new, const double* shift,
                                const double* fbt, double* scal, long long* nodes, int W, double dt, double sqrt_dt,
                                double h0p_re, double h0p_im, double w_floor, double w_cap, double damping,
                                hipStream_t s) {
  if (W <= 0) return hipSuccess;
  hipLaunchKernelGGL(weight_update_kernel, dim3(1), dim3(1024), 0, s, weights, ov_old, ov_new, shift, fbt, scal, nodes,
                     W, dt, sqrt_dt, h0p_re, h0p_im, w_floor, w_cap, damping);
  return hipGetLastError();
}

hipError_t launch_chol_rinv(const double* gram, double* rinv, int n, int W, hipStream_t s, double* dprod,
                            int accumulate, int* flags) {
  if (W <= 0) return hipSuccess;
  if (n < 1 || n > TROT_MAX_NOCC) return hipErrorInvalidValue;
  const size_t lds = (size_t)2 * n * n * 16;
  hipError_t e = set_lds((const void*)chol_rinv_kernel, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(chol_rinv_kernel, dim3(W), dim3(256), lds, s, gram, rinv, n, dprod, accumulate, flags);
  return hipGetLastError();
}

hipError_t launch_qr_fallback(const double* src, double* q, const int* flags, int N, int n, int W, double* rinv,
                              double* dprod, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  if (n < 1 || n > N) return hipErrorInvalidValue;
  const size_t lds = (size_t)n * 16;
  hipLaunchKernelGGL(qr_fallback_kernel, dim3(W), dim3(256), lds, s, src, q, flags, N, n, rinv, dprod);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gauss-Jordan inverse and determinant for n > 64 (inv_det_kernel covers n <= 64 in registers): partial
// pivoting on |re| + |im| with the first index on ties (LAPACK's izamax), rows swapped in place, the inverse
// built in place and its columns swapped back at the end; det = prod(pivots) (-1)^swaps.  One workgroup of
// 256 threads per matrix.  The matrix lives in LDS when it fits (n <= INV_LDS_MAX), else in place in the
// workgroup's own n x n block of minv (m and minv must not overlap).  Fixed reduction order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void inv_det_big_kernel(const double* __restrict__ m, double* __restrict__ minv,
                                                          double* __restrict__ det, int n) {
  extern __shared__ double sh[];
  const long w = blockIdx.x;
  const int t = threadIdx.x;
  const long nn = (long)n * n;
  const bool in_lds = n <= INV_LDS_MAX;
  char* base = reinterpret_cast<char*>(sh) + (in_lds ? nn * 16 : 0);
  cplx* a = in_lds ? reinterpret_cast<cplx*>(sh) : reinterpret_cast<cplx*>(minv + 2 * w * nn);
  cplx* colk = reinterpret_cast<cplx*>(base);
  double* rv = reinterpret_cast<double*>(base + (long)n * 16);
  int* ri = reinterpret_cast<int*>(base + (long)n * 16 + 256 * 8);
  int* ipiv = ri + 256;
  const cplx* mw = reinterpret_cast<const cplx*>(m + 2 * w * nn);
  for (long e = t; e < nn; e += 256) a[e] = mw[e];
  __syncthreads();
  cplx d = {1.0, 0.0};  // thread 0
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = n;
    for (int i = k + t; i < n; i += 256) {
      const double v = cabs1(a[(long)i * n + k]);
      if (v > best) {
        best = v;
        bi = i;
      }
    }
    rv[t] = best;
    ri[t] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (t < s) {
        const double v2 = rv[t + s];
        const int i2 = ri[t + s];
        if (v2 > rv[t] || (v2 == rv[t] && i2 < ri[t])) {
          rv[t] = v2;
          ri[t] = i2;
        }
      }
      __syncthreads();
    }
    int piv_row = ri[0];
    if (piv_row >= n) piv_row = k;  // no comparable entry (NaN column): keep the row
    __syncthreads();
    if (t == 0) ipiv[k] = piv_row;
    if (piv_row != k)
      for (int j = t; j < n; j += 256) {
        const cplx x = a[(long)k * n + j];
        a[(long)k * n + j] = a[(long)piv_row * n + j];
        a[(long)piv_row * n + j] = x;
      }
    __syncthreads();
    const cplx piv = a[(long)k * n + k];
    if (t == 0) {
      d = cmul(d, piv);
      if (piv_row != k) d = {-d.re, -d.im};
    }
    const cplx ip = cinv(piv);
    for (int i = t; i < n; i += 256) colk[i] = i == k ? cplx{0.0, 0.0} : a[(long)i * n + k];
    __syncthreads();
    for (int j = t; j < n; j += 256) a[(long)k * n + j] = cmul(j == k ? cplx{1.0, 0.0} : a[(long)k * n + j], ip);
    __syncthreads();
    for (long e = t; e < nn; e += 256) {
      const int i = (int)(e / n), j = (int)(e - (long)i * n);
      if (i == k) continue;
      const cplx x = j == k ? cplx{0.0, 0.0} : a[e];
      a[e] = csub(x, cmul(colk[i], a[(long)k * n + j]));
    }
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; --k) {
    const int q = ipiv[k];
    if (q != k)
      for (int i = t; i < n; i += 256) {
        const cplx x = a[(long)i * n + k];
        a[(long)i * n + k] = a[(long)i * n + q];
        a[(long)i * n + q] = x;
      }
    __syncthreads();
  }
  if (in_lds) {
    cplx* o = reinterpret_cast<cplx*>(minv + 2 * w * nn);
    for (long e = t; e < nn; e += 256) o[e] = a[e];
  }
  if (t == 0) cstore(det, w, d);
}

// x[i] = a value in [-1, 1) hashed from i (deterministic, non-zero operands for timing).
__global__ __launch_bounds__(256) void fill_pattern_kernel(double* __restrict__ x, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) {
    unsigned long long h = (unsigned long long)i * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 32;
    x[i] = (double)(h >> 11) * (1.0 / 4503599627370496.0) - 1.0;  // 2^-52 (h >> 11) - 1
  }
}

hipError_t launch_fill_pattern(double* x, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_pattern_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, n);
  return hipGetLastError();
}

hipError_t launch_sr_index(const double* weights, double* cw, int* idx, double* w_new, int W, double zeta,
                           const double* zeta_dev, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  hipLaunchKernelGGL(sr_index_kernel, dim3(1), dim3(1024), 0, s, weights, cw, idx, w_new, W, zeta, zeta_dev);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const double* src, double* dst, const int* idx, int W, long row, hipStream_t s) {
  if (W <= 0 || row <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(W), dim3(256), 0, s, src, dst, idx, row);
  return hipGetLastError();
}

hipError_t launch_mf_trace(const double* chol, const double* dm, int G, int N, double* mf, hipStream_t s) {
  if (G <= 0) return hipSuccess;
  hipLaunchKernelGGL(mf_trace_kernel, dim3(G), dim3(256), 0, s, chol, dm, N, mf);
  return hipGetLastError();
}

hipError_t launch_shift_grouped(const ShiftProb* probs, const int* start, int nprob, int total_w, hipStream_t s) {
  if (nprob <= 0 || total_w <= 0) return hipSuccess;
  hipLaunchKernelGGL(shift_grouped_kernel, dim3(total_w), dim3(256), 0, s, probs, start, nprob);
  return hipGetLastError();
}

hipError_t launch_sum_grouped(const SumProb* probs, const int* start, int nprob, int total_blocks, hipStream_t s) {
  if (nprob <= 0 || total_blocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(sum_grouped_kernel, dim3(total_blocks), dim3(256), 0, s, probs, start, nprob);
  return hipGetLastError();
}

hipError_t launch_update_grouped(const UpdProb* probs, int nprob, hipStream_t s) {
  if (nprob <= 0) return hipSuccess;
  hipLaunchKernelGGL(update_grouped_kernel, dim3(nprob), dim3(1024), 0, s, probs);
  return hipGetLastError();
}

hipError_t launch_sum_batches(const double* src, double* dst, int nb, long len, double scale, int accumulate,
                              hipStream_t s) {
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(sum_batches_kernel, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, src, dst, nb, len,
                     scale, accumulate);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// block measurement tail (trot/prop/blocks.py:236-257), single workgroup:
//   e = Re E_L; non-finite or |e - E_est| > sqrt(2/dt) -> E_est; non-finite
//   walkers get weight 0; e_block = sum w e / sum w (E_est if sum w == 0);
//   E_est <- (1 - alpha) E_est + alpha e_block.
// scal[1] = e_estimate (in/out); out[0] = e_block, out[1] = sum w.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void block_energy_kernel(const double* __restrict__ energy,
                                                            double* __restrict__ weights, double* __restrict__ scal,
                                                            double* __restrict__ out, int W, double thresh,
                                                            double alpha) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  const double e_ref = scal[1];
  double sw = 0.0, swe = 0.0;
  for (int w = t; w < W; w += 1024) {
    double e = energy[2 * w];
    const bool bad = !isfinite(e);
    if (bad || fabs(e - e_ref) > thresh) e = e_ref;
    double wt = weights[w];
    if (bad) {
      wt = 0.0;
      weights[w] = 0.0;
    }
    sw += wt;
    swe += wt * e;
  }
  sw = block_sum<1024>(sw, red);
  swe = block_sum<1024>(swe, red);
  if (t == 0) {
    const double e_block = (sw == 0.0) ? e_ref : swe / sw;
    scal[1] = (1.0 - alpha) * e_ref + alpha * e_block;
    out[0] = e_block;
    out[1] = sw;
  }
}

hipError_t launch_block_energy(const double* energy, double* weights, double* scal, double* out, int W,
                               double thresh, double alpha, hipStream_t s) {
  if (W <= 0) return hipSuccess;
  hipLaunchKernelGGL(block_energy_kernel, dim3(1), dim3(1024), 0, s, energy, weights, scal, out, W, thresh, alpha);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// chol (G, N, N) -> sym-block storage (G, P*64) of the upper-triangle 8 x 8
// blocks (common.hpp), FP64 or FP32.  Grid (nb, nb, G), 64 threads; blocks
// with bi > bj exit.  Entries outside N x N are zero.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void pack_symblk_kernel(const double* __restrict__ chol, T* __restrict__ out, int N,
                                                         int nb) {
  const int bi = blockIdx.x, bj = blockIdx.y;
  if (bi > bj) return;
  const long g = blockIdx.z;
  const int t = threadIdx.x;
  const int i = bi * 8 + (t >> 3), j = bj * 8 + (t & 7);
  const double v = (i < N && j < N) ? chol[g * (long)N * N + (long)i * N + j] : 0.0;
  out[g * symblk_count(nb) * 64 + symblk_index(bi, bj, nb) * 64 + t] = (T)v;
}

hipError_t launch_pack_symblk(const double* chol, void* out, int out_f32, int G, int N, hipStream_t s) {
  if (G <= 0 || N <= 0) return hipSuccess;
  const int nb = (N + 7) / 8;
  dim3 grid(nb, nb, G);
  if (G > 65535) return hipErrorInvalidConfiguration;
  if (out_f32)
    hipLaunchKernelGGL(pack_symblk_kernel<float>, grid, dim3(64), 0, s, chol, (float*)out, N, nb);
  else
    hipLaunchKernelGGL(pack_symblk_kernel<double>, grid, dim3(64), 0, s, chol, (double*)out, N, nb);
  return hipGetLastError();
}

// complex128 -> complex64 (n complex elements)
__global__ void c128_to_c64_kernel(const double* __restrict__ src, float* __restrict__ dst, long n) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double2 v = reinterpret_cast<const double2*>(src)[e];
  reinterpret_cast<float2*>(dst)[e] = make_float2((float)v.x, (float)v.y);
}

hipError_t launch_c128_to_c64(const double* src, float* dst, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(c128_to_c64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}
